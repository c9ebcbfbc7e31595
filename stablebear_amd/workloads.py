"""Seeded synthetic PCF workloads of the benchmark configurations.

Each generator returns CSR arrays ``(offsets[int64 n+1], t[float64], v[float64])``.
Recipe: ``numpy.random.default_rng(seed)``; record counts m drawn first for all
PCFs, then per PCF ``t = [0, cumsum(Exponential(scale, m - 1))]`` and the values.

* ``c1``: N=1000, m ~ U{10..100}, scale 1, values m-1, m-2, ..., 0 (seed 1001).
* ``c2``: N=20000, m ~ U{50..500}, scale 1, values N(0,1) with the last set to 0
  (seed 2002).
* ``c3``: 256 problems, problem b seeded 3003*1000+b: N_b ~ U{64..4096},
  m ~ round(LogNormal(4, 1)) clipped to [5, 20000], scale 1, integer values in
  [-50, 50] with the last set to 0.
* ``c4``: N=4096, m ~ U{20..200}, scale 1, N(0,1) values ending at 0 (seed 4004);
  weights W_ij ~ U(0,1) for i < j (symmetric).
* ``c5``: X and Y with N=50000 each (seeds [5005, 0] and [5005, 1]),
  m ~ U{500..5000}, increments Exponential with rate 0.1 (scale 10), N(0,1)
  values ending at 0.

Smaller ``n`` gives the same recipe at a reduced size (parity-test cases).
"""

from __future__ import annotations

import numpy as np


def _times(rng, m, scale):
    M = int(m.sum())
    inc = rng.exponential(scale, size=M)
    off = np.zeros(len(m) + 1, dtype=np.int64)
    np.cumsum(m, out=off[1:])
    t = np.empty(M, dtype=np.float64)
    for i in range(len(m)):
        b, e = off[i], off[i + 1]
        seg = inc[b:e].copy()
        seg[0] = 0.0
        np.cumsum(seg, out=t[b:e])
    return off, t


def _normal_values(rng, off):
    v = rng.standard_normal(int(off[-1]))
    v[off[1:] - 1] = 0.0
    return v


def c1(n=1000, seed=1001):
    rng = np.random.default_rng(seed)
    m = rng.integers(10, 101, size=n).astype(np.int64)
    off, t = _times(rng, m, 1.0)
    v = np.empty_like(t)
    for i in range(n):
        b, e = off[i], off[i + 1]
        v[b:e] = np.arange(e - b - 1, -1, -1, dtype=np.float64)
    return off, t, v


def c2(n=20000, seed=2002):
    rng = np.random.default_rng(seed)
    m = rng.integers(50, 501, size=n).astype(np.int64)
    off, t = _times(rng, m, 1.0)
    return off, t, _normal_values(rng, off)


def c3_problem(b, n=None):
    rng = np.random.default_rng(3003 * 1000 + b)
    nb = int(rng.integers(64, 4097)) if n is None else int(n)
    m = np.clip(np.round(rng.lognormal(4.0, 1.0, size=nb)), 5, 20000).astype(np.int64)
    off, t = _times(rng, m, 1.0)
    v = rng.integers(-50, 51, size=int(off[-1])).astype(np.float64)
    v[off[1:] - 1] = 0.0
    return off, t, v


def c3(k=256, n=None):
    """List of k problems (each an (offsets, t, v) triple)."""
    return [c3_problem(b, n) for b in range(k)]


def c4(n=4096, seed=4004):
    rng = np.random.default_rng(seed)
    m = rng.integers(20, 201, size=n).astype(np.int64)
    off, t = _times(rng, m, 1.0)
    v = _normal_values(rng, off)
    w = rng.uniform(0.0, 1.0, size=n * (n - 1) // 2)
    return off, t, v, w


def c5_side(n=50000, seed=5005, stream=0, m_lo=500, m_hi=5000):
    """One side (X: stream 0, Y: stream 1) of c5 as a CSR triple."""
    rng = np.random.default_rng([seed, stream])
    m = rng.integers(m_lo, m_hi + 1, size=n).astype(np.int64)
    off, t = _times(rng, m, 10.0)
    return off, t, _normal_values(rng, off)


def c5(n=50000, seed=5005, m_lo=500, m_hi=5000):
    """(X, Y) CSR triples."""
    return (c5_side(n, seed, 0, m_lo, m_hi), c5_side(n, seed, 1, m_lo, m_hi))


def to_pcfs(off, t, v, dtype=np.float64):
    """List of Pcf objects of a CSR triple."""
    from .pcf import Pcf

    pts = np.stack([t, v], axis=1).astype(dtype)
    return [Pcf(pts[off[i]:off[i + 1]]) for i in range(len(off) - 1)]
