"""GPU tests of the hand-written backward kernel behind torch.autograd:
gradients of sum(W * D) (L2 pdist, p = 2) and sum(W * K) (l2_kernel) with
respect to values and breakpoint times, against the CPU analytic gradient of
the oracle and against central finite differences (h = 1e-6, FP64)."""

import numpy as np
import pytest
import torch

import oracle
from stablebear_amd import autograd, workloads

pytestmark = pytest.mark.gpu


def _inputs(n, seed=4004):
    off, t, v, w = workloads.c4(n, seed=seed)
    return off, t, v, w


def _grads(kind, prec, off, t, v, W):
    vt = torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True)
    tt = torch.tensor(t, dtype=torch.float64, device="cuda", requires_grad=True)
    fn = autograd.pdist_l2 if kind == "pdist" else autograd.l2_kernel
    out = fn(vt, tt, off, precision=prec)
    Wt = torch.tensor(W, dtype=out.dtype, device="cuda")
    loss = (Wt * out).sum()
    loss.backward()
    return out.detach().double().cpu().numpy(), vt.grad.cpu().numpy(), tt.grad.cpu().numpy()


def _interior_times(off):
    first = np.zeros(off[-1], dtype=bool)
    first[off[:-1]] = True
    return ~first


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("kind", ["pdist", "kernel"])
def test_gradient_matches_cpu_analytic(gpu, kind, prec):
    off, t, v, w = _inputs(300)
    n = len(off) - 1
    if kind == "kernel":
        W = np.random.default_rng(1).uniform(0, 1, n * (n + 1) // 2)
        op, okind = oracle.OP_DOT, oracle.KIND_KERNEL
    else:
        W = w
        op, okind = oracle.OP_LP, oracle.KIND_PDIST
    out, gv, gt = _grads(kind, prec, off, t, v, W)
    fwd = oracle.pairwise(op, 2.0 if op == oracle.OP_LP else 1.0, okind, off, t, v)
    ev, et = oracle.grad(op, okind, off, t, v, W, fwd if op == oracle.OP_LP else None)
    tol = 1e-9 if prec == "fp64" else 2e-5
    scale_v = np.abs(ev).max()
    scale_t = np.abs(et).max()
    assert np.max(np.abs(gv - ev)) <= tol * scale_v
    m = _interior_times(off)
    assert np.max(np.abs(gt[m] - et[m])) <= tol * scale_t
    assert np.all(gt[~m] == 0.0)


@pytest.mark.parametrize("kind", ["pdist", "kernel"])
def test_gradient_with_times_shared_between_pcfs(gpu, kind):
    """Breakpoint times on an integer grid, so that many PCFs (in one band
    and across bands) share times: the backward sees the ties between PCFs
    and takes the mean of the limits there, as the oracle does. (Sets whose
    PCFs share no time take the walk without tie flags, checked by the
    tests above.)"""
    rng = np.random.default_rng(77)
    n = 200
    m = rng.integers(3, 25, size=n)
    off = np.concatenate([[0], np.cumsum(m)]).astype(np.int64)
    t = np.concatenate([np.concatenate([[0.0], np.sort(rng.choice(np.arange(1, 60), k - 1, replace=False))])
                        for k in m]).astype(np.float64)
    v = rng.standard_normal(off[-1])
    v[off[1:] - 1] = 0.0
    if kind == "kernel":
        W = rng.uniform(0, 1, n * (n + 1) // 2)
        op, okind = oracle.OP_DOT, oracle.KIND_KERNEL
    else:
        W = rng.uniform(0, 1, n * (n - 1) // 2)
        op, okind = oracle.OP_LP, oracle.KIND_PDIST
    out, gv, gt = _grads(kind, "fp64", off, t, v, W)
    fwd = oracle.pairwise(op, 2.0 if op == oracle.OP_LP else 1.0, okind, off, t, v)
    ev, et = oracle.grad(op, okind, off, t, v, W, fwd if op == oracle.OP_LP else None)
    assert np.max(np.abs(gv - ev)) <= 1e-9 * np.abs(ev).max()
    mi = _interior_times(off)
    assert np.max(np.abs(gt[mi] - et[mi])) <= 1e-9 * np.abs(et).max()


def test_gradient_matches_finite_differences(gpu):
    off, t, v, w = _inputs(40, seed=7)
    n = len(off) - 1
    _, gv, gt = _grads("pdist", "fp64", off, t, v, w)

    def loss(tt, vv):
        return float(np.dot(w, oracle.pairwise(oracle.OP_LP, 2.0, 0, off, tt, vv, nthreads=1)))

    rng = np.random.default_rng(2)
    last = set((off[1:] - 1).tolist())
    h = 1e-6
    for k in [k for k in rng.choice(len(v), 25, replace=False) if k not in last]:
        vp, vm = v.copy(), v.copy()
        vp[k] += h
        vm[k] -= h
        fd = (loss(t, vp) - loss(t, vm)) / (2 * h)
        assert gv[k] == pytest.approx(fd, rel=1e-5, abs=1e-6)
    interior = np.nonzero(_interior_times(off))[0]
    for k in rng.choice(interior, 25, replace=False):
        others = np.delete(t, k)
        if np.min(np.abs(others - t[k])) < 1e-4:
            continue
        tp, tm = t.copy(), t.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss(tp, v) - loss(tm, v)) / (2 * h)
        assert gt[k] == pytest.approx(fd, rel=1e-5, abs=1e-6)
    del n


def test_c4_full_size_gradient(gpu):
    """C4 at full size (N = 4096): FP32C forward, backward vs the CPU analytic gradient."""
    off, t, v, w = _inputs(4096)
    out, gv, gt = _grads("pdist", "fp32c", off, t, v, w)
    fwd = oracle.pairwise(oracle.OP_LP, 2.0, 0, off, t, v)
    assert np.max(np.abs(out - fwd) / fwd) <= 1e-6
    ev, et = oracle.grad(oracle.OP_LP, oracle.KIND_PDIST, off, t, v, w, fwd)
    assert np.max(np.abs(gv - ev)) <= 2e-5 * np.abs(ev).max()
    m = _interior_times(off)
    assert np.max(np.abs(gt[m] - et[m])) <= 2e-5 * np.abs(et).max()


def test_c4_full_size_kernel_gradient(gpu):
    """C4 l2_kernel at full size (N = 4096): FP32C forward against the FP64
    oracle (normwise: the Gram entries are signed sums) and backward of
    sum(W * K) against the CPU analytic gradient."""
    off, t, v, _ = _inputs(4096)
    n = len(off) - 1
    W = np.random.default_rng(11).uniform(0, 1, n * (n + 1) // 2)
    out, gv, gt = _grads("kernel", "fp32c", off, t, v, W)
    fwd = oracle.pairwise(oracle.OP_DOT, 1.0, oracle.KIND_KERNEL, off, t, v)
    scale = oracle.pairwise(oracle.OP_DOT, 1.0, oracle.KIND_KERNEL, off, t, np.abs(v))
    assert np.max(np.abs(out - fwd) / scale) <= 1e-6
    ev, et = oracle.grad(oracle.OP_DOT, oracle.KIND_KERNEL, off, t, v, W, None)
    assert np.max(np.abs(gv - ev)) <= 2e-5 * np.abs(ev).max()
    m = _interior_times(off)
    assert np.max(np.abs(gt[m] - et[m])) <= 2e-5 * np.abs(et).max()


def test_c4_full_size_kernel_finite_differences(gpu):
    """Central finite differences (h = 1e-6) of sum(W * K) at N = 4096, the
    loss evaluated by the FP64 GPU forward, against the FP64 backward, on
    sampled value and interior-time coordinates away from coincident
    breakpoints."""
    off, t, v, _ = _inputs(4096)
    n = len(off) - 1
    W = np.random.default_rng(12).uniform(0, 1, n * (n + 1) // 2)
    _, gv, gt = _grads("kernel", "fp64", off, t, v, W)
    Wd = torch.tensor(W, dtype=torch.float64, device="cuda")
    off_d = torch.tensor(off, dtype=torch.int64, device="cuda")
    owner = np.repeat(np.arange(n), np.diff(off))

    def loss(tt, vv, k):
        # only the entries K_ij of the PCF i owning point k change: the loss
        # difference is taken over them, so the rounding of the other
        # 8.4e6 terms does not enter the difference quotient
        i = int(owner[k])
        j = np.arange(n, dtype=np.int64)
        idx = np.where(j <= i, i * (i + 1) // 2 + j, j * (j + 1) // 2 + i)
        idx_d = torch.from_numpy(idx).cuda()
        with torch.no_grad():
            out = autograd.l2_kernel(torch.tensor(vv, dtype=torch.float64, device="cuda"),
                                     torch.tensor(tt, dtype=torch.float64, device="cuda"),
                                     off_d, precision="fp64")
            return float(torch.dot(Wd[idx_d], out[idx_d]).item())

    rng = np.random.default_rng(13)
    last = set((off[1:] - 1).tolist())
    h = 1e-6
    for k in [k for k in rng.choice(len(v), 12, replace=False) if k not in last]:
        vp, vm = v.copy(), v.copy()
        vp[k] += h
        vm[k] -= h
        fd = (loss(t, vp, k) - loss(t, vm, k)) / (2 * h)
        assert gv[k] == pytest.approx(fd, rel=1e-5, abs=1e-6)
    ts = np.sort(t)
    interior = np.nonzero(_interior_times(off))[0]
    checked = 0
    for k in rng.choice(interior, 40, replace=False):
        pos = np.searchsorted(ts, t[k])
        gap = min(ts[pos] - ts[pos - 1] if pos > 0 else np.inf,
                  ts[pos + 1] - ts[pos] if pos + 1 < len(ts) else np.inf)
        if gap < 1e-4:
            continue
        tp, tm = t.copy(), t.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss(tp, v, k) - loss(tm, v, k)) / (2 * h)
        assert gt[k] == pytest.approx(fd, rel=1e-5, abs=1e-6)
        checked += 1
        if checked == 12:
            break
    assert checked >= 6


def test_cached_set_keeps_each_forwards_values(gpu):
    """The set and plan of one (offsets, times) pair are cached and their
    values refreshed on the device: two forwards with different values share
    them, and each backward still differentiates its own forward."""
    autograd.clear_cache()
    off, t, v, w = _inputs(120, seed=21)
    v2 = v[::-1].copy()
    v2[off[1:] - 1] = 0.0
    tt = torch.tensor(t, dtype=torch.float64, device="cuda", requires_grad=True)
    va = torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True)
    vb = torch.tensor(v2, dtype=torch.float64, device="cuda", requires_grad=True)
    off_d = torch.tensor(off, dtype=torch.int64, device="cuda")
    W = torch.tensor(w, dtype=torch.float64, device="cuda")
    da = autograd.pdist_l2(va, tt, off_d, precision="fp64")
    db = autograd.pdist_l2(vb, tt, off_d, precision="fp64")
    assert len(autograd._cache) == 1 and len(next(iter(autograd._cache.values()))) == 1
    ((W * da).sum() + 2.0 * (W * db).sum()).backward()
    for vv, scale, d, g in ((v, 1.0, da, va.grad), (v2, 2.0, db, vb.grad)):
        fwd = oracle.pairwise(oracle.OP_LP, 2.0, 0, off, t, vv)
        np.testing.assert_allclose(d.detach().cpu().numpy(), fwd, rtol=1e-13)
        ev, _ = oracle.grad(oracle.OP_LP, oracle.KIND_PDIST, off, t, vv, scale * w, fwd)
        assert np.max(np.abs(g.cpu().numpy() - ev)) <= 1e-9 * np.abs(ev).max()
    _, et_a = oracle.grad(oracle.OP_LP, oracle.KIND_PDIST, off, t, v, w,
                          oracle.pairwise(oracle.OP_LP, 2.0, 0, off, t, v))
    _, et_b = oracle.grad(oracle.OP_LP, oracle.KIND_PDIST, off, t, v2, 2.0 * w,
                          oracle.pairwise(oracle.OP_LP, 2.0, 0, off, t, v2))
    m = _interior_times(off)
    et = et_a + et_b
    assert np.max(np.abs(tt.grad.cpu().numpy()[m] - et[m])) <= 1e-9 * np.abs(et).max()


@pytest.mark.parametrize("op", ["lp", "dot"])
def test_repeated_backward_reuses_sorted_lists(gpu, op):
    """A plan's first backward sorts the set's breakpoints into event lists;
    later backwards keep that order and refresh only the value jumps. After
    the set's values change on the device, the second backward of the same
    plan matches a backward over a new set with those values (sorted
    afresh) and the oracle's analytic gradient."""
    from stablebear_amd import runtime

    off, t, v, w = _inputs(500, seed=41)
    v2 = np.random.default_rng(42).standard_normal(len(v))
    v2[off[1:] - 1] = 0.0
    s = torch.cuda.current_stream().cuda_stream
    p = 2.0 if op == "lp" else 1.0

    def run(plan, npts):
        out = torch.empty(plan.out_count, dtype=torch.float64, device="cuda")
        plan.execute(out.data_ptr(), True, stream=s)
        g = torch.from_numpy(np.random.default_rng(43).uniform(0, 1, plan.out_count)).cuda()
        gv = torch.full((npts,), np.nan, dtype=torch.float64, device="cuda")
        gt = torch.full((npts,), np.nan, dtype=torch.float64, device="cuda")
        plan.backward(g.data_ptr(), out.data_ptr(), True, gv.data_ptr(), gt.data_ptr(), True, stream=s)
        torch.cuda.synchronize()
        return g.cpu().numpy(), out.cpu().numpy(), gv.cpu().numpy(), gt.cpu().numpy()

    ds = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp64")
    plan = runtime.Plan(ds, op=op, p=p)
    run(plan, len(v))  # sorts the lists
    v2_d = torch.tensor(v2, dtype=torch.float64, device="cuda")
    ds.update_values(v2_d.data_ptr(), s)
    g, fwd, gv, gt = run(plan, len(v))  # reuses them
    ds_new = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v2), "fp64")
    _, _, gv_new, gt_new = run(runtime.Plan(ds_new, op=op, p=p), len(v))
    assert np.max(np.abs(gv - gv_new)) <= 1e-12 * np.abs(gv_new).max()
    assert np.max(np.abs(gt - gt_new)) <= 1e-12 * np.abs(gt_new).max()
    oop = oracle.OP_LP if op == "lp" else oracle.OP_DOT
    okind = oracle.KIND_PDIST if op == "lp" else oracle.KIND_KERNEL
    ev, et = oracle.grad(oop, okind, off, t, v2, g, fwd if op == "lp" else None)
    assert np.max(np.abs(gv - ev)) <= 1e-9 * np.abs(ev).max()
    m = _interior_times(off)
    assert np.max(np.abs(gt[m] - et[m])) <= 1e-9 * np.abs(et).max()
