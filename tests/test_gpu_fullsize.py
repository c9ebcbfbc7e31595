"""GPU parity at the benchmark sizes, through checks that do not need the
oracle over the whole output: sampled rows against the oracle (bit-exact in
FP64), FP32C against FP64 on every entry (on the device), the polarization
identity D^2 = K_ii + K_jj - 2 K_ij between the L2 distance and the L2 kernel,
and a batched launch against per-problem launches."""

import numpy as np
import pytest
import torch

import oracle
from stablebear_amd import runtime, workloads
from stablebear_amd.batched import BatchedPlan

pytestmark = pytest.mark.gpu


def _exec(plan, f64):
    out = torch.empty(plan.out_count, dtype=torch.float64 if f64 else torch.float32, device="cuda")
    plan.execute(out.data_ptr(), f64, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return out


def _row(off, t, v, i, op, p):
    """Oracle entries (i, j), j > i, of a pdist (as a 1 x (n-i-1) cdist)."""
    n = len(off) - 1
    a, b = off[i], off[i + 1]
    fx = (np.array([0, b - a]), t[a:b], v[a:b])
    c = off[i + 1]
    gy = (off[i + 1:] - c, t[c:], v[c:])
    return oracle.pairwise(op, p, oracle.KIND_CDIST, *fx, *gy)[: n - i - 1]


def _row_idx(n, i):
    j = np.arange(i + 1, n, dtype=np.int64)
    return j * (j - 1) // 2 + i


def _max_rel_dev(a32, b64):
    a = a32.double()
    nz = b64 != 0
    return float(((a[nz] - b64[nz]).abs() / b64[nz].abs()).max().item())


def test_c2_full_l1_pdist(gpu):
    n = 20000
    off, t, v = workloads.c2(n)
    csr = runtime.csr_from_arrays(off, t, v)
    out64 = _exec(runtime.Plan(runtime.DeviceSet(csr, "fp64"), op="l1"), True)
    out32 = _exec(runtime.Plan(runtime.DeviceSet(csr, "fp32c"), op="l1"), False)
    assert _max_rel_dev(out32, out64) <= 1e-6
    assert bool((out64 > 0).all().item())
    for i in (0, 1, 7777, n - 2):
        idx = torch.from_numpy(_row_idx(n, i)).cuda()
        np.testing.assert_array_equal(out64[idx].cpu().numpy(), _row(off, t, v, i, oracle.OP_L1, 1.0))


def test_c2_full_l2_and_kernel_polarization(gpu):
    n = 20000
    off, t, v = workloads.c2(n)
    csr = runtime.csr_from_arrays(off, t, v)
    ds = runtime.DeviceSet(csr, "fp64")
    D = _exec(runtime.Plan(ds, op="lp", p=2.0), True)
    K = _exec(runtime.Plan(ds, op="dot"), True)
    rng = np.random.default_rng(5)
    i = rng.integers(0, n, 20000)
    j = rng.integers(0, n, 20000)
    keep = i != j
    i, j = np.maximum(i[keep], j[keep]), np.minimum(i[keep], j[keep])
    d = D[torch.from_numpy(i * (i - 1) // 2 + j).cuda()].cpu().numpy()
    kij = K[torch.from_numpy(i * (i + 1) // 2 + j).cuda()].cpu().numpy()
    kii = K[torch.from_numpy(i * (i + 1) // 2 + i).cuda()].cpu().numpy()
    kjj = K[torch.from_numpy(j * (j + 1) // 2 + j).cuda()].cpu().numpy()
    assert np.max(np.abs(d * d - (kii + kjj - 2 * kij)) / (kii + kjj)) <= 1e-12
    for r in (3, 12000):
        idx = torch.from_numpy(_row_idx(n, r)).cuda()
        np.testing.assert_allclose(D[idx].cpu().numpy(), _row(off, t, v, r, oracle.OP_LP, 2.0),
                                   rtol=1e-13, atol=0)


def test_c3_full_batched_against_single_problems(gpu):
    probs = workloads.c3()
    b32 = BatchedPlan(probs, op="l1", precision="fp32c")
    b64 = BatchedPlan(probs, op="l1", precision="fp64")
    out32 = torch.empty(b32.out_count, dtype=torch.float32, device="cuda")
    out64 = torch.empty(b64.out_count, dtype=torch.float64, device="cuda")
    g = b32.graph(out32.data_ptr(), False)
    g.launch(torch.cuda.current_stream().cuda_stream)
    b64.execute(out64.data_ptr(), True, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert _max_rel_dev(out32, out64) <= 1e-6
    # FP64 batched == FP64 single-problem launches == oracle, on a spread of problems
    for b in (0, 17, 101, 255):
        off, t, v = probs[b]
        single = _exec(runtime.Plan(runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp64"),
                                    op="l1"), True)
        a, e = b64.offsets[b], b64.offsets[b] + single.numel()
        assert torch.equal(out64[a:e], single)
        if (len(off) - 1) <= 700:
            np.testing.assert_array_equal(single.cpu().numpy(),
                                          oracle.pairwise(oracle.OP_L1, 1.0, 0, off, t, v))


def test_c5_shape_cdist_rows_and_means(gpu):
    """The 50k x 50k cdist's per-pair shape (m ~ U{500..5000}, long pairs with
    the FP64 fold) at 4000 x 4000."""
    X, Y = workloads.c5(4000)
    sx = runtime.DeviceSet(runtime.csr_from_arrays(*X), "fp32c")
    sy = runtime.DeviceSet(runtime.csr_from_arrays(*Y), "fp32c")
    plan = runtime.Plan(sx, sy, op="l1")
    assert not plan.long_pairs
    out = torch.empty(plan.out_count, dtype=torch.float32, device="cuda")
    rs = torch.empty(4000, dtype=torch.float64, device="cuda")
    plan.execute(out.data_ptr(), False, rs.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    D = out.view(4000, 4000)
    # row sums add the FP64 per-pair results; D holds them rounded to FP32
    # (<= 6e-8 relative each), so the two sums agree to that
    np.testing.assert_allclose(rs.cpu().numpy(), D.double().sum(dim=1).cpu().numpy(), rtol=1e-7)
    xo, xt, xv = X
    for i in (0, 1999, 3999):
        a, b = xo[i], xo[i + 1]
        exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_CDIST,
                              np.array([0, b - a]), xt[a:b], xv[a:b], *Y)
        got = D[i].double().cpu().numpy()
        assert np.max(np.abs(got - exp) / exp) <= 1e-6


def test_c5_full_cdist_rows_past_2_31_and_means(gpu):
    """The full 50,000 x 50,000 L1 cdist (m ~ U{500..5000}, FP32C with the
    FP64 fold, 2.5e9 outputs, so output indices pass 2^31): fused row means
    against the means of the written rows of D, and whole rows at the start
    and at the end of the index range against the reference walk."""
    n = 50000
    X, Y = workloads.c5(n)
    sx = runtime.DeviceSet(runtime.csr_from_arrays(*X), "fp32c", pack="gpu")
    sy = runtime.DeviceSet(runtime.csr_from_arrays(*Y), "fp32c", pack="gpu")
    plan = runtime.Plan(sx, sy, op="l1")
    assert plan.out_count == n * n > 2 ** 31
    out = torch.empty(plan.out_count, dtype=torch.float32, device="cuda")
    rs = torch.empty(n, dtype=torch.float64, device="cuda")
    plan.execute(out.data_ptr(), False, rs.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    D = out.view(n, n)
    means = torch.empty(n, dtype=torch.float64, device="cuda")
    for r0 in range(0, n, 5000):
        means[r0:r0 + 5000] = D[r0:r0 + 5000].double().mean(dim=1)
    np.testing.assert_allclose(rs.cpu().numpy() / n, means.cpu().numpy(), rtol=1e-7)
    assert bool((D[-1] > 0).all().item()) and bool(torch.isfinite(D[-3:]).all().item())
    xo, xt, xv = X
    for i in (0, 1, 42950, n - 2, n - 1):  # row 42950 starts past index 2^31
        a, b = xo[i], xo[i + 1]
        exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_CDIST,
                              np.array([0, b - a]), xt[a:b], xv[a:b], *Y)
        got = D[i].double().cpu().numpy()
        assert np.max(np.abs(got - exp) / exp) <= 1e-6, i
