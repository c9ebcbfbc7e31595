"""Write tests/golden/ref_vectors.npz: small seeded inputs and the outputs of the
reference's own compiled walk (oracle/_ref/libsbref.so, built from the
reference headers by oracle/Makefile).

Run in a container that has the reference tree: ``make -C oracle && python
tests/golden/make_ref_fixtures.py``. The .npz holds data only (inputs and
expected outputs); it lets parity tests run where the reference is absent.
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from stablebear_amd import workloads  # noqa: E402


def _edge_set():
    """Duplicate times, cross-PCF ties, constant and zero PCFs, negative values."""
    pcfs = [
        [[0, 3], [1, 2], [4, 5], [6, 0]],
        [[0, 2], [3, 4], [4, 2], [5, 1], [8, 3]],
        [[0, 0]],
        [[0, 7]],
        [[0, 1], [0, 4], [2, 2], [2, 5], [2, -1], [4, 0]],
        [[0, -3], [1, 6], [4, 2], [6, 0]],
        [[0, 2.5], [1, 2.5], [3, 0], [8, 0]],
        [[0, 1], [4, 0]],
        [[0, 0], [0.5, 1.25], [1.75, -0.5], [3.0, 0.0]],
    ]
    off = np.zeros(len(pcfs) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(p) for p in pcfs])
    pts = np.concatenate([np.asarray(p, dtype=np.float64) for p in pcfs])
    return off, pts[:, 0].copy(), pts[:, 1].copy()


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libsbref.so is not built (needs the reference tree)")
    out = {}
    cases = {
        "c1": workloads.c1(60),
        "c2": workloads.c2(40),
        "edge": _edge_set(),
        "c4": workloads.c4(30)[:3],
    }
    ops = {"l1": (oracle.OP_L1, 1.0), "l2": (oracle.OP_LP, 2.0), "l3": (oracle.OP_LP, 3.0),
           "dot": (oracle.OP_DOT, 1.0)}
    for name, (off, t, v) in cases.items():
        out[f"{name}_off"], out[f"{name}_t"], out[f"{name}_v"] = off, t, v
        for oname, (op, p) in ops.items():
            kind = oracle.KIND_KERNEL if op == oracle.OP_DOT else oracle.KIND_PDIST
            out[f"{name}_{oname}_f64"] = oracle.pairwise(op, p, kind, off, t, v, use_ref=True)
            out[f"{name}_{oname}_f32"] = oracle.pairwise(op, p, kind, off, t, v,
                                                         dtype=np.float32, use_ref=True)
    X, Y = workloads.c5(12, m_lo=20, m_hi=60)
    for key, (off, t, v) in (("cx", X), ("cy", Y)):
        out[f"{key}_off"], out[f"{key}_t"], out[f"{key}_v"] = off, t, v
    for oname, (op, p) in (("l1", (oracle.OP_L1, 1.0)), ("l2", (oracle.OP_LP, 2.0))):
        out[f"cdist_{oname}_f64"] = oracle.pairwise(op, p, oracle.KIND_CDIST, *X, *Y,
                                                    use_ref=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
