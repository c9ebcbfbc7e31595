"""GPU tests of the row reductions (``sbamd_rows_reduce`` / ``sbamd_rows_knn``)
against numpy on the same matrices: dense and compressed distance matrices,
FP32 and FP64, ties (integer-valued entries: the lower column first), rows
shorter than k (padding), k from 1 to 32. Sums are FP64 in a different order
than numpy's, so they are held to 1e-12 relative; minima, maxima and the
k nearest (values and columns) are exact."""

import numpy as np
import pytest
import torch

import stablebear_amd as sb
from stablebear_amd import rows

pytestmark = pytest.mark.gpu


def _ref_knn(dense, k, skip_diag):
    nr, nc = dense.shape
    vals = np.full((nr, k), np.inf, dtype=dense.dtype)
    idx = np.full((nr, k), -1, dtype=np.int64)
    for i in range(nr):
        cols = np.array([j for j in range(nc) if not (skip_diag and j == i)], dtype=np.int64)
        if cols.size == 0:
            continue
        row = dense[i, cols]
        order = np.lexsort((cols, row))[:k]  # by value, then column
        vals[i, : order.size] = row[order]
        idx[i, : order.size] = cols[order]
    return vals, idx


def _ref_stats(dense, skip_diag):
    d = dense.astype(np.float64)
    if skip_diag:
        mask = ~np.eye(d.shape[0], dtype=bool)
        s = np.array([d[i, mask[i]].sum() for i in range(d.shape[0])])
        lo = np.array([d[i, mask[i]].min() if d.shape[0] > 1 else np.inf for i in range(d.shape[0])])
        hi = np.array([d[i, mask[i]].max() if d.shape[0] > 1 else -np.inf for i in range(d.shape[0])])
        return s, lo, hi
    return d.sum(axis=1), d.min(axis=1), d.max(axis=1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1), (5, 3), (37, 300), (64, 1000)])
def test_dense_rows(gpu, dtype, shape):
    rng = np.random.default_rng(shape[0] * 31 + shape[1])
    dense = rng.integers(0, 20, shape).astype(dtype)  # many ties
    s, mean, lo, hi = rows.row_stats(dense)
    rs, rlo, rhi = _ref_stats(dense, False)
    np.testing.assert_allclose(s, rs, rtol=1e-12)
    np.testing.assert_allclose(mean, rs / shape[1], rtol=1e-12)
    np.testing.assert_array_equal(lo, rlo)
    np.testing.assert_array_equal(hi, rhi)
    for k in (1, 3, 8, 16, 32):
        v, i = rows.knn(dense, k)
        rv, ri = _ref_knn(dense, k, False)
        np.testing.assert_array_equal(v, rv)
        np.testing.assert_array_equal(i, ri)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 9, 150, 700])
def test_compressed_distance_matrix_rows(gpu, dtype, n):
    rng = np.random.default_rng(n)
    storage = rng.random(n * (n - 1) // 2).astype(dtype)
    if storage.size:
        storage[::7] = storage[0]  # ties
    D = sb.DistanceMatrix._from_storage(n, storage)
    dense = D.to_dense()
    s, mean, lo, hi = rows.row_stats(D)
    rs, rlo, rhi = _ref_stats(dense, True)
    np.testing.assert_allclose(s, rs, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(lo, rlo)
    np.testing.assert_array_equal(hi, rhi)
    for k in (1, 4, 9, 32):
        v, i = rows.knn(D, k)
        rv, ri = _ref_knn(dense, k, True)
        np.testing.assert_array_equal(v, rv)
        np.testing.assert_array_equal(i, ri)


def test_device_inputs_stay_on_the_device(gpu):
    from stablebear_amd import workloads

    off, t, v = workloads.c1(120)
    D = sb.pdist(workloads.to_pcfs(off, t, v))
    dev = torch.from_numpy(D.storage.copy()).cuda()
    vals, idx = rows.knn(dev, 5, n=120)
    assert vals.is_cuda and idx.is_cuda
    hv, hi = rows.knn(D, 5)
    np.testing.assert_array_equal(vals.cpu().numpy(), hv)
    np.testing.assert_array_equal(idx.cpu().numpy(), hi)
    m = rows.row_means(dev, n=120)
    np.testing.assert_allclose(m.cpu().numpy(), rows.row_means(D), rtol=0, atol=0)
    with pytest.raises(ValueError, match="needs n"):
        rows.knn(dev, 3)
    with pytest.raises(ValueError, match="k must be"):
        rows.knn(D, 33)
