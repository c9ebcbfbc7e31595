"""The allocation-free backward (sbamd_plan_backward_noalloc) against the
allocating one (sbamd_plan_backward), through the C-ABI: two calls in a row
on one GPU, the second with a smaller set, for the L2 pdist and the L2
kernel. The two run the same kernels; the FP64 accumulators are added with
atomics, so their order (and last bits) may differ between runs: the bar is
1e-12 of the largest gradient entry."""

import numpy as np
import pytest
import torch

from stablebear_amd import runtime, workloads

pytestmark = pytest.mark.gpu


def _plan(n, seed, op):
    off, t, v, _ = workloads.c4(n, seed=seed)
    ds = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp32c")
    plan = runtime.Plan(ds, op=op, p=2.0 if op == "lp" else 1.0)
    return ds, plan, len(v)


def _both(plan, npts, seed):
    s = torch.cuda.current_stream().cuda_stream
    out = torch.empty(plan.out_count, dtype=torch.float32, device="cuda")
    plan.execute(out.data_ptr(), False, stream=s)
    g = torch.from_numpy(np.random.default_rng(seed).uniform(0, 1, plan.out_count)).float().cuda()
    res = []
    for fn in (plan.backward, plan.backward_noalloc):
        gv = torch.full((npts,), np.nan, dtype=torch.float64, device="cuda")
        gt = torch.full((npts,), np.nan, dtype=torch.float64, device="cuda")
        fn(g.data_ptr(), out.data_ptr(), False, gv.data_ptr(), gt.data_ptr(), True, stream=s)
        torch.cuda.synchronize()
        res.append((gv.cpu().numpy(), gt.cpu().numpy()))
    return res


@pytest.mark.parametrize("op", ["lp", "dot"])
def test_noalloc_backward_matches_allocating_one(gpu, op):
    keep = []
    for n, seed in ((2000, 31), (300, 32)):  # the second call on a smaller set
        ds, plan, npts = _plan(n, seed, op)
        keep.append((ds, plan))
        (av, at), (bv, bt) = _both(plan, npts, seed)
        assert np.isfinite(av).all() and np.isfinite(bv).all()
        assert np.isfinite(at).all() and np.isfinite(bt).all()
        assert np.max(np.abs(bv - av)) <= 1e-12 * np.abs(av).max()
        assert np.max(np.abs(bt - at)) <= 1e-12 * np.abs(at).max()


def test_noalloc_backward_rejects_sets_beyond_its_memory(gpu):
    rng = np.random.default_rng(33)
    m = np.full(2700, 200, dtype=np.int64)  # 540,000 records > 2^19
    off, t = workloads._times(rng, m, 1.0)
    v = workloads._normal_values(rng, off)
    ds = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp32c")
    plan = runtime.Plan(ds, op="lp", p=2.0)
    g = torch.zeros(plan.out_count, dtype=torch.float32, device="cuda")
    out = torch.zeros(plan.out_count, dtype=torch.float32, device="cuda")
    gv = torch.zeros(len(v), dtype=torch.float64, device="cuda")
    gt = torch.zeros(len(v), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="working memory holds at most"):
        plan.backward_noalloc(g.data_ptr(), out.data_ptr(), False, gv.data_ptr(), gt.data_ptr(), True)
