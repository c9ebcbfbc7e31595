"""GPU tests of device-side set packing (``sbamd_set_create_gpu``): the packed
layout (records, record offsets, permutation, source points) equals the host
packer's byte for byte for SoA and AoS host views and for device views with
any strides; the breakpoint rules raise the host packer's errors with the
same messages; and a pdist over a device-packed set is bit-exact with the
oracle (FP64) as the host-packed set is."""

import numpy as np
import pytest
import torch

import oracle
from stablebear_amd import _lib as L
from stablebear_amd import runtime

pytestmark = pytest.mark.gpu


def _random_set(rng, n, dtype, max_m=300):
    """CSR set with empty, single-point, duplicate-time and multi-chunk PCFs."""
    lens = rng.integers(0, max_m, n)
    lens[: min(n, 4)] = [0, 1, 64, 65][: min(n, 4)]
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    t = np.empty(off[-1], dtype=dtype)
    v = rng.standard_normal(off[-1]).astype(dtype)
    for i in range(n):
        b, e = off[i], off[i + 1]
        if e == b:
            continue
        inc = rng.exponential(1.0, e - b - 1)
        inc[rng.random(e - b - 1) < 0.1] = 0.0  # duplicate times
        t[b:e] = np.concatenate([[0.0], np.cumsum(inc)]).astype(dtype)
        v[e - 1] = 0.0
    return off, t, v


def _layout(ds):
    rec, roff, perm, src = ds.export()
    return rec, roff, perm, src


def _assert_same_layout(a, b):
    for x, y, name in zip(_layout(a), _layout(b), ("records", "roff", "perm", "src")):
        np.testing.assert_array_equal(x, y, err_msg=name)


@pytest.mark.parametrize("dtype,prec", [(np.float64, "fp64"), (np.float64, "fp32c"),
                                        (np.float32, "fp32c")])
def test_soa_host_view_matches_host_packer(gpu, dtype, prec):
    rng = np.random.default_rng(11)
    off, t, v = _random_set(rng, 300, dtype)
    csr = runtime.csr_from_arrays(off, t, v)
    _assert_same_layout(runtime.DeviceSet(csr, prec), runtime.DeviceSet(csr, prec, pack="gpu"))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_aos_host_view_matches_host_packer(gpu, dtype):
    rng = np.random.default_rng(12)
    off, t, v = _random_set(rng, 200, dtype)
    pts = np.ascontiguousarray(np.stack([t, v], axis=1))  # AoS {t, v}
    es = pts.itemsize
    view = L.PcfView()
    view.n = len(off) - 1
    view.offsets = off.ctypes.data
    view.t = pts.ctypes.data
    view.v = pts.ctypes.data + es
    view.t_stride = view.v_stride = 2 * es
    view.dtype = L.F64 if dtype == np.float64 else L.F32
    view.location = L.HOST
    gpu_set = runtime.DeviceSet(None, "fp32c", view=view, keep=(off, pts))
    host_set = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp32c")
    _assert_same_layout(host_set, gpu_set)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_device_views_read_in_place(gpu, dtype):
    rng = np.random.default_rng(13)
    np_dt = np.float64 if dtype == torch.float64 else np.float32
    off, t, v = _random_set(rng, 257, np_dt)
    host_set = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp32c")
    d_off = torch.from_numpy(off).cuda()
    # contiguous arrays
    ds = runtime.DeviceSet.from_device(d_off, torch.from_numpy(t).cuda(),
                                       torch.from_numpy(v).cuda(), "fp32c")
    _assert_same_layout(host_set, ds)
    # columns of an (M, 2) point tensor: element stride 2
    pts = torch.from_numpy(np.stack([t, v], axis=1)).cuda()
    ds2 = runtime.DeviceSet.from_device(d_off, pts[:, 0], pts[:, 1], "fp32c")
    _assert_same_layout(host_set, ds2)


def test_large_pcfs_and_empty_sets(gpu):
    rng = np.random.default_rng(14)
    # PCFs far longer than a wavefront, and an all-empty set
    off, t, v = _random_set(rng, 20, np.float64, max_m=5000)
    csr = runtime.csr_from_arrays(off, t, v)
    _assert_same_layout(runtime.DeviceSet(csr, "fp64"), runtime.DeviceSet(csr, "fp64", pack="gpu"))
    z = runtime.csr_from_arrays(np.zeros(4, dtype=np.int64), np.zeros(0), np.zeros(0))
    _assert_same_layout(runtime.DeviceSet(z, "fp64"), runtime.DeviceSet(z, "fp64", pack="gpu"))
    e = runtime.csr_from_arrays(np.zeros(1, dtype=np.int64), np.zeros(0), np.zeros(0))
    ds = runtime.DeviceSet(e, "fp64", pack="gpu")
    assert ds.n == 0 and ds.records == 0


def _bad_sets():
    base_t = np.array([0.0, 1.0, 2.0, 3.0, 0.0, 0.5])
    base_v = np.array([1.0, 2.0, 3.0, 0.0, 1.0, 0.0])
    off = np.array([0, 4, 6], dtype=np.int64)
    dec = base_t.copy()
    dec[2] = 0.5  # decreasing in PCF 0
    nonzero = base_t.copy()
    nonzero[4] = 0.25  # PCF 1 starts at 0.25
    nan = base_t.copy()
    nan[5] = np.nan
    inf = base_t.copy()
    inf[3] = np.inf
    return [(off, dec, base_v), (off, nonzero, base_v), (off, nan, base_v), (off, inf, base_v)]


@pytest.mark.parametrize("case", range(4))
def test_rule_violations_raise_the_host_packers_errors(gpu, case):
    off, t, v = _bad_sets()[case]
    csr = runtime.csr_from_arrays(off, t, v)
    with pytest.raises(ValueError) as host_err:
        runtime.DeviceSet(csr, "fp64")
    with pytest.raises(ValueError) as gpu_err:
        runtime.DeviceSet(csr, "fp64", pack="gpu")
    assert str(gpu_err.value) == str(host_err.value)


def test_bad_offsets_raise(gpu):
    t = np.zeros(3)
    csr = runtime.csr_from_arrays(np.array([0, 2, 1, 3]), t, t)
    with pytest.raises(ValueError, match="non-decreasing"):
        runtime.DeviceSet(csr, "fp64", pack="gpu")
    csr = runtime.csr_from_arrays(np.array([1, 2, 3]), t, t)
    with pytest.raises(ValueError, match="start at 0"):
        runtime.DeviceSet(csr, "fp64", pack="gpu")


def test_pdist_over_a_device_packed_set_is_bit_exact(gpu):
    rng = np.random.default_rng(15)
    off, t, v = _random_set(rng, 150, np.float64)
    ds = runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp64", pack="gpu")
    got = runtime.Plan(ds, op="l1").execute_to_host(True)
    exp = oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST, off, t, v)
    np.testing.assert_array_equal(got, exp)
