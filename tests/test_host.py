"""Host-side logic (no GPU): the C-ABI library loads and exports every declared
symbol, input validation, segment counting, tile planning coverage, and the
Python containers / PCF / tensor semantics of the drop-in API."""

import os
import re

import numpy as np
import pytest

import stablebear_amd as sb
from stablebear_amd import _lib as L
from stablebear_amd import runtime
from stablebear_amd.distance_matrix import DistanceMatrix
from stablebear_amd.symmetric_matrix import SymmetricMatrix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "stablebear_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sbamd_\w+)\s*\(", hdr, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = _declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(L.EXPORTED) == names


def test_version_and_device_count():
    assert b"gfx950" in L.lib().sbamd_version()
    assert L.device_count() >= 0


def _csr(pcfs, dtype=np.float64):
    return runtime.csr_from_arrays(
        np.concatenate([[0], np.cumsum([len(p) for p in pcfs])]),
        np.concatenate([np.asarray(p, dtype)[:, 0] for p in pcfs]) if pcfs else np.zeros(0, dtype),
        np.concatenate([np.asarray(p, dtype)[:, 1] for p in pcfs]) if pcfs else np.zeros(0, dtype))


def _validate(csr):
    s = csr.struct()
    import ctypes

    return L.check(L.lib().sbamd_validate_set(ctypes.byref(s)))


def test_validation_rules():
    _validate(_csr([[[0, 1], [1, 2], [1, 3], [4, 0]], [[0, 5]]]))
    with pytest.raises(ValueError, match="non-decreasing"):
        _validate(_csr([[[0, 1], [2, 2], [1, 3]]]))
    with pytest.raises(ValueError, match="t=0"):
        _validate(_csr([[[0.5, 1], [2, 2]]]))
    with pytest.raises(ValueError, match="finite"):
        _validate(_csr([[[0, 1], [np.inf, 2]]]))
    with pytest.raises(ValueError, match="finite"):
        _validate(_csr([[[0, 1], [np.nan, 2]]]))
    bad = _csr([[[0, 1]]])
    bad.offsets[1] = -1
    with pytest.raises(ValueError):
        _validate(bad)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_validation_of_large_sets_reports_the_first_offending_pcf(dtype):
    """Sets of 2^20 points or more are scanned by several threads; the
    error is still the one of the lowest offending PCF, for each rule."""
    n, m = 3000, 500  # 1.5e6 points
    off = np.arange(n + 1, dtype=np.int64) * m
    t = np.tile(np.arange(m, dtype=dtype), n)
    v = np.ones(n * m, dtype=dtype)
    _validate(runtime.csr_from_arrays(off, t, v))
    cases = [("non-decreasing", lambda t, i: t.__setitem__(off[i] + 7, t[off[i] + 5])),
             ("t=0", lambda t, i: t.__setitem__(off[i], 0.5)),
             ("finite", lambda t, i: t.__setitem__(off[i] + 9, np.inf)),
             ("finite", lambda t, i: t.__setitem__(off[i] + 9, np.nan))]
    for (first_msg, first), (_, later) in zip(cases, cases[1:] + cases[:1]):
        tt = t.copy()
        later(tt, 2900)  # in the last thread's range
        first(tt, 1700)
        later(tt, 2100)
        with pytest.raises(ValueError, match=first_msg):
            _validate(runtime.csr_from_arrays(off, tt, v))


def test_count_segments_formula_and_brute_force():
    rng = np.random.default_rng(4)
    pcfs = []
    for m in rng.integers(1, 9, size=13):
        t = np.concatenate([[0.0], np.sort(rng.uniform(0, 10, size=m - 1))])
        pcfs.append(np.stack([t, rng.standard_normal(m)], 1))
    # a duplicate time: counted once after zero-length pieces are dropped
    pcfs.append(np.array([[0.0, 1.0], [2.0, 3.0], [2.0, 4.0], [5.0, 0.0]]))
    csr = _csr(pcfs)
    m = np.array([len(np.unique(p[:, 0])) for p in pcfs])
    n = len(pcfs)
    pd = sum(m[i] + m[j] - 1 for i in range(n) for j in range(i + 1, n))
    assert runtime.count_segments(csr) == pd
    kd = sum(m[i] + m[j] - 1 for i in range(n) for j in range(i + 1, n)) + m.sum()
    assert runtime.count_segments(csr, include_diagonal=True) == kd
    cd = sum(m[i] + m[j] - 1 for i in range(n) for j in range(n))
    assert runtime.count_segments(csr, csr) == cd


def _tiles(m_rows, m_cols, R, C):
    import ctypes

    mr = np.asarray(m_rows, dtype=np.int64)
    mc = None if m_cols is None else np.asarray(m_cols, dtype=np.int64)
    cnt = ctypes.c_int64()
    perm = np.zeros(len(mr), dtype=np.int64)
    L.check(L.lib().sbamd_plan_tiles_host(len(mr), mr.ctypes.data_as(L.c_i64p),
                                          0 if mc is None else len(mc),
                                          None if mc is None else mc.ctypes.data_as(L.c_i64p),
                                          R, C, None, 0, ctypes.byref(cnt),
                                          perm.ctypes.data_as(L.c_i64p)))
    tiles = np.zeros((cnt.value, 4), dtype=np.int32)
    L.check(L.lib().sbamd_plan_tiles_host(len(mr), mr.ctypes.data_as(L.c_i64p),
                                          0 if mc is None else len(mc),
                                          None if mc is None else mc.ctypes.data_as(L.c_i64p),
                                          R, C, tiles.ctypes.data_as(L.c_i32p), cnt.value,
                                          ctypes.byref(cnt), None))
    return tiles, perm


@pytest.mark.parametrize("n,R,C", [(1, 16, 16), (2, 16, 16), (17, 16, 16), (33, 8, 8), (100, 16, 16),
                                   (100, 32, 16), (100, 16, 32), (70, 24, 16), (33, 32, 16), (2, 32, 16),
                                   (47, 8, 32)])
def test_symmetric_tiles_cover_lower_triangle_once(n, R, C):
    """Square or rectangular tiles of the strict lower triangle: every wanted
    pair in exactly one tile; a tile without the diagonal flag holds only
    wanted pairs, and every flagged tile holds an unwanted one."""
    m = np.random.default_rng(n).integers(1, 50, size=n)
    tiles, perm = _tiles(m, None, R, C)
    assert np.all(np.diff(m[perm]) >= 0)  # sorted by record count
    seen = np.zeros((n, n), dtype=int)
    for r0, c0, nr, ncf in tiles:
        nc, diag = ncf & 0xFFFF, bool(ncf >> 16)
        assert 1 <= nr <= R and 1 <= nc <= C and r0 % R == 0 and c0 % C == 0
        a, b = np.meshgrid(np.arange(r0, r0 + nr), np.arange(c0, c0 + nc), indexing="ij")
        wanted = a > b
        assert wanted.any()
        assert diag == (not wanted.all())
        seen[a[wanted], b[wanted]] += 1
    assert np.array_equal(seen, np.tril(np.ones((n, n), dtype=int), -1))


@pytest.mark.parametrize("nx,ny,R,C", [(5, 7, 16, 16), (40, 3, 8, 16), (33, 65, 16, 8)])
def test_cross_tiles_cover_grid_once(nx, ny, R, C):
    rng = np.random.default_rng(nx * ny)
    tiles, _ = _tiles(rng.integers(1, 50, size=nx), rng.integers(1, 50, size=ny), R, C)
    seen = np.zeros((nx, ny), dtype=int)
    for r0, c0, nr, nc in tiles:
        seen[r0:r0 + nr, c0:c0 + nc] += 1
    assert np.all(seen == 1)


@pytest.mark.parametrize("nprob,seed", [(1, 1), (3, 2), (40, 3)])
def test_xcd_queues_partition_the_tile_list(nprob, seed):
    """The per-XCD queues of a batched plan are a permutation of its tile list
    cut into 8 contiguous queues: a supertile unit (G x G tiles of one
    problem) stays in one queue, its tiles in their given order, and the
    queues' work is balanced to within the largest unit."""
    rng = np.random.default_rng(seed)
    R = 16
    tiles, work = [], []
    for b in range(nprob):
        n = int(rng.integers(20, 700))
        nt = (n + R - 1) // R
        for i in range(nt):
            for j in range(i + 1):
                tiles.append((i * R, j * R, b))
                work.append(int(rng.integers(1, 1000)))
    order_in = np.argsort(-np.array(work), kind="stable")  # longest first, as the planner gives them
    t = np.array([tiles[k] for k in order_in], dtype=np.int32)
    w = np.array([work[k] for k in order_in], dtype=np.int64)
    order = np.zeros(len(t), dtype=np.int64)
    qoff = np.zeros(9, dtype=np.int64)
    L.check(L.lib().sbamd_queue_order_host(len(t), t.ctypes.data_as(L.c_i32p), w.ctypes.data_as(L.c_i64p),
                                           R, R, 8, order.ctypes.data_as(L.c_i64p),
                                           qoff.ctypes.data_as(L.c_i64p)))
    assert np.array_equal(np.sort(order), np.arange(len(t)))
    assert qoff[0] == 0 and qoff[8] == len(t) and np.all(np.diff(qoff) >= 0)
    queue_of = np.searchsorted(qoff, np.arange(len(t)), side="right") - 1
    units = {}
    nunits = {G: len({(p, r // (G * R), c // (G * R)) for r, c, p in t}) for G in (16, 8, 4, 2, 1)}
    G = next(g for g in (16, 8, 4, 2, 1) if nunits[g] >= 64 or g == 1)
    for pos, k in enumerate(order):
        r, c, p = t[k]
        units.setdefault((p, r // (G * R), c // (G * R)), []).append((queue_of[pos], k))
    umax = 0
    for members in units.values():
        assert len({q for q, _ in members}) == 1
        ks = [k for _, k in members]
        assert ks == sorted(ks)  # the unit's tiles keep their order
        umax = max(umax, int(w[ks].sum()))
    loads = [int(w[order[qoff[x]:qoff[x + 1]]].sum()) for x in range(8)]
    assert max(loads) - min(loads) <= umax


# ------------------------------------------------------- containers -------

@pytest.mark.parametrize("dtype", [sb.float32, sb.float64])
def test_distance_matrix_semantics(dtype):
    dm = DistanceMatrix(5, dtype=dtype)
    assert dm.size == 5 and dm.storage_count == 10
    dm[0, 3] = 7.0
    assert dm[3, 0] == 7.0 and dm[-2, 0] == 7.0 and dm[1, 1] == 0.0
    with pytest.raises(ValueError):
        dm[0, 1] = -1.0
    with pytest.raises(ValueError):
        dm[1, 1] = 5.0
    dm[1, 1] = 0.0
    with pytest.raises(IndexError):
        _ = dm[5, 0]
    assert DistanceMatrix(0, dtype=dtype).to_dense().shape == (0, 0)
    assert DistanceMatrix(1, dtype=dtype).storage_count == 0
    with pytest.raises(TypeError, match="Unsupported dtype"):
        DistanceMatrix(3, dtype=int)
    d = dm.to_dense()
    assert d.dtype == (np.float32 if dtype is sb.float32 else np.float64)
    assert np.array_equal(DistanceMatrix.from_dense(d).to_dense(), d)
    assert np.array_equal(np.asarray(dm), d)
    import pickle

    assert pickle.loads(pickle.dumps(dm)) == dm


def test_distance_matrix_from_dense_validation():
    with pytest.raises(ValueError, match="Diagonal"):
        DistanceMatrix.from_dense(np.array([[1.0, 0.0], [0.0, 0.0]]))
    with pytest.raises(ValueError, match="nonnegative"):
        DistanceMatrix.from_dense(np.array([[0.0, -1.0], [-1.0, 0.0]]))
    with pytest.raises(ValueError, match="symmetric"):
        DistanceMatrix.from_dense(np.array([[0.0, 1.0], [2.0, 0.0]]))
    with pytest.raises(TypeError):
        DistanceMatrix.from_dense(np.array([[0, 1], [1, 0]], dtype=np.int32))
    dm = DistanceMatrix.from_dense(np.array([[0.0, 1.0, 2.0], [1.0, 0.0, 3.0], [2.0, 3.0, 0.0]]))
    assert (dm[0, 1], dm[0, 2], dm[1, 2]) == (1.0, 2.0, 3.0)
    assert list(dm.storage) == [1.0, 2.0, 3.0]  # index max*(max-1)/2 + min


def test_symmetric_matrix_semantics():
    sm = SymmetricMatrix(4)
    assert sm.storage_count == 10
    sm[2, 1] = -3.0
    sm[3, 3] = 5.0
    assert sm[1, 2] == -3.0 and sm[-1, -1] == 5.0
    d = sm.to_dense()
    assert np.array_equal(SymmetricMatrix.from_dense(d).to_dense(), d)
    with pytest.raises(ValueError, match="symmetric"):
        SymmetricMatrix.from_dense(np.array([[0.0, 1.0], [2.0, 0.0]]))
    s3 = SymmetricMatrix.from_dense(np.arange(9.0).reshape(3, 3) + np.arange(9.0).reshape(3, 3).T)
    assert list(s3.storage) == [0.0, 4.0, 8.0, 8.0, 12.0, 16.0]  # index max*(max+1)/2 + min


# ------------------------------------------------------ PCF / tensor ------

def test_pcf_construction_rules():
    f = sb.Pcf(np.array([[0.0, 1.0], [1.0, 2.0], [3.0, 0.0]], dtype=np.float32))
    assert f.size == 3 and f.dtype is sb.pcf32
    assert sb.Pcf(np.zeros((0, 2))).to_numpy().tolist() == [[0.0, 0.0]]
    with pytest.raises(ValueError, match="non-decreasing"):
        sb.Pcf(np.array([[0.0, 1.0], [2.0, 1.0], [1.0, 0.0]]))
    with pytest.raises(ValueError, match="t=0"):
        sb.Pcf(np.array([[1.0, 1.0]]))
    with pytest.raises(RuntimeError, match="Nx2"):
        sb.Pcf(np.zeros((3, 3)))
    assert sb.Pcf([[0, 4], [1, 9]]).dtype is sb.pcf64
    assert sb.Pcf(np.array([[0, 1]], dtype=np.int32)).dtype is sb.pcf32i
    g = sb.Pcf(np.array([[0.0, 4.0], [1.0, 9.0]]))
    assert g(0.5) == 4.0 and g(1.0) == 9.0


def test_tensor_views_and_assignment():
    X = sb.zeros((4, 6), dtype=sb.pcf64)
    for i in range(4):
        for j in range(6):
            X[i, j] = sb.Pcf(np.array([[0.0, float(i * 6 + j + 1)], [1.0, 0.0]]))
    V = X[::2, 1::2]
    assert V.shape == (2, 3)
    assert [p.to_numpy()[0, 1] for p in V.flat_pcfs()] == [2.0, 4.0, 6.0, 14.0, 16.0, 18.0]
    R = X[:, 2][::-1]
    assert [p.to_numpy()[0, 1] for p in R.flat_pcfs()] == [21.0, 15.0, 9.0, 3.0]
    with pytest.raises(TypeError):
        X[0, 0] = sb.Pcf(np.array([[0.0, 1.0]], dtype=np.float32))
    assert sb.zeros((3,)).dtype is sb.pcf32


def test_csr_packing_offsets():
    pcfs = [sb.Pcf(np.array([[0.0, 1.0], [1.0, 2.0], [2.0, 3.0]])), sb.Pcf(np.array([[0.0, 4.0]])),
            sb.Pcf(np.array([[0.0, 5.0], [3.0, 6.0]]))]
    csr = runtime.csr_from_pcfs(pcfs, sb.pcf64)
    assert csr.offsets.tolist() == [0, 3, 4, 6]  # test_offset_data_manager.cpp:23-55
    assert csr.v.tolist() == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]


def test_api_validation_without_device():
    with pytest.raises(ValueError, match="p must be >= 1"):
        sb.pdist(sb.zeros((2,)), p=0.5)
    with pytest.raises(ValueError, match="p must be >= 1"):
        sb.cdist(sb.zeros((2,)), sb.zeros((3,)), p=0)
    for p in (0.5, 0, -1):  # test_norms.py:8-23
        with pytest.raises(ValueError, match="p must be >= 1"):
            sb.lp_norm(sb.zeros((2,)), p=p)
    with pytest.raises(ValueError, match="1d tensor"):
        sb.pdist(sb.zeros((10, 20)))
    with pytest.raises(ValueError, match="1d tensor"):
        sb.l2_kernel(sb.zeros((10, 20)))
    with pytest.raises(TypeError):
        sb.lp_distance(sb.Pcf(np.array([[0.0, 1.0]])), 42)
    with pytest.raises(TypeError):
        sb.lp_distance(sb.Pcf(np.array([[0.0, 1.0]], dtype=np.float32)),
                       sb.Pcf(np.array([[0.0, 1.0]])))
    with pytest.raises(TypeError):
        sb.lp_distance(sb.Pcf(np.array([[0, 1], [1, 0]], dtype=np.int32)),
                       sb.Pcf(np.array([[0, 0]], dtype=np.int32)))
    with pytest.raises(TypeError, match="same dtype"):
        sb.cdist(sb.PcfTensor([sb.Pcf(np.array([[0.0, 1.0]], dtype=np.float32))]),
                 sb.PcfTensor([sb.Pcf(np.array([[0.0, 1.0]]))]))
    # empty and single inputs need no device
    assert sb.pdist(sb.zeros((0,))).size == 0
    assert sb.pdist([]).size == 0
    D = sb.pdist(sb.zeros((1,)))
    assert D.size == 1 and D[0, 0] == 0.0
    assert sb.l2_kernel(sb.zeros((0,))).size == 0
    assert sb.cdist(sb.zeros((0,)), sb.zeros((3,))).shape == (0, 3)


def test_system_knobs():
    with pytest.raises(RuntimeError):
        sb.system.force_cpu(True)
    sb.system.force_cpu(False)
    sb.system.set_cuda_threshold(0)
    with pytest.raises(ValueError):
        sb.system.set_block_size(40, 40)
    with pytest.raises(ValueError):
        sb.system.set_tile_cap(1)
    with pytest.raises(ValueError):
        sb.system.set_output_budget(-1)
    sb.system.set_precision("fp32c")
    assert sb.system.get_precision() == "fp32c"
    sb.system.set_precision("fp64")
    assert sb.system.build_type() == "HIP"
