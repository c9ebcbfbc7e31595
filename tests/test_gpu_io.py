"""GPU tests of loading a file of PCFs straight into a device set
(``sbamd_set_create_image``): the packed layout equals the host packer's for
the same PCFs (pcf32 / pcf64, 1-D and N-D tensors in row-major order, and a
single PCF object), a pdist over it is bit-exact with the oracle, rule
violations in the file raise the host packer's errors, and files that do not
hold floating-point PCFs are rejected."""

import numpy as np
import pytest

import oracle
import stablebear_amd as sb
from stablebear_amd import runtime, workloads

pytestmark = pytest.mark.gpu


def _tensor(off, t, v, np_dt, shape=None):
    pcfs = [sb.Pcf(np.stack([t[off[i]:off[i + 1]], v[off[i]:off[i + 1]]], 1).astype(np_dt))
            for i in range(len(off) - 1)]
    arr = np.empty(len(pcfs), dtype=object)
    arr[:] = pcfs
    return sb.PcfTensor(arr.reshape(shape or (len(pcfs),)))


def _same(a, b):
    for x, y, name in zip(a.export(), b.export(), ("records", "roff", "perm", "src")):
        np.testing.assert_array_equal(x, y, err_msg=name)


@pytest.mark.parametrize("np_dt,prec", [(np.float64, "fp64"), (np.float64, "fp32c"),
                                        (np.float32, "fp32c")])
def test_file_set_matches_host_packer(gpu, tmp_path, np_dt, prec):
    off, t, v = workloads.c2(120, seed=7)
    t, v = t.astype(np_dt), v.astype(np_dt)
    path = tmp_path / "x.sb"
    sb.save(_tensor(off, t, v, np_dt, (10, 12)), str(path))
    ds = sb.io.load_device_set(str(path), prec)
    _same(runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), prec), ds)


def test_file_set_pdist_bit_exact_and_object_file(gpu, tmp_path):
    off, t, v = workloads.c1(200)
    path = tmp_path / "c1.sb"
    sb.save(_tensor(off, t, v, np.float64), str(path))
    ds = sb.io.load_device_set(str(path), "fp64")
    got = runtime.Plan(ds, op="l1").execute_to_host(True)
    np.testing.assert_array_equal(got, oracle.pairwise(oracle.OP_L1, 1.0, oracle.KIND_PDIST,
                                                       off, t, v))
    f = sb.Pcf(np.array([[0.0, 1.0], [1.0, 2.0], [3.0, 0.0]]))
    sb.save(f, str(tmp_path / "f.sb"))
    one = sb.io.load_device_set(str(tmp_path / "f.sb"), "fp64")
    assert one.n == 1 and one.records == 3


def test_file_rule_violations_and_foreign_files(gpu, tmp_path):
    # the second PCF's times decrease (0, 2, 1): the host packer's error
    off = np.array([0, 2, 5], dtype=np.int64)
    t = np.array([0.0, 1.0, 0.0, 2.0, 1.0])
    v = np.array([1.0, 0.0, 1.0, 2.0, 0.0])
    valid = sb.io.dumps(sb.zeros((2,), dtype=sb.pcf64))
    head = valid[: sb.io._info(np.frombuffer(valid, dtype=np.uint8)).payload]
    body = b"".join(np.uint64(off[i + 1] - off[i]).tobytes()
                    + np.stack([t[off[i]:off[i + 1]], v[off[i]:off[i + 1]]], 1).tobytes()
                    for i in range(2))
    path = tmp_path / "bad.sb"
    path.write_bytes(head + body)
    with pytest.raises(ValueError) as gpu_err:
        sb.io.load_device_set(str(path), "fp64")
    with pytest.raises(ValueError) as host_err:
        runtime.DeviceSet(runtime.csr_from_arrays(off, t, v), "fp64")
    assert str(gpu_err.value) == str(host_err.value)
    sb.save(np.zeros(3), str(tmp_path / "num.sb"))
    with pytest.raises(ValueError, match="does not hold PCFs"):
        sb.io.load_device_set(str(tmp_path / "num.sb"))
