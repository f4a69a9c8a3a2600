"""Seeded synthetic Max-Cut instances for the five workloads, and the
objective-matrix construction of process_input.c:235-299.

Edge lists are 1-based (rudy format, process_input.c:197-225).
"""
from __future__ import annotations

import numpy as np


def objective_matrix(nv, ei, ej, w):
    """SP->L: Laplacian of the graph on vertices 0..nv-2 in the upper-left
    block, its row sums in the last row/column, their total in the corner
    (process_input.c:235-299).  A repeated edge overwrites."""
    A = np.zeros((nv, nv))
    A[np.asarray(ei) - 1, np.asarray(ej) - 1] = w
    A[np.asarray(ej) - 1, np.asarray(ei) - 1] = w
    deg = A.sum(axis=1)
    Lap = np.diag(deg) - A
    L = np.zeros((nv, nv))
    L[: nv - 1, : nv - 1] = Lap[: nv - 1, : nv - 1]
    rs = L[: nv - 1, : nv - 1].sum(axis=1)
    L[nv - 1, : nv - 1] = rs
    L[: nv - 1, nv - 1] = rs
    L[nv - 1, nv - 1] = rs.sum()
    return np.asfortranarray(L)


def read_rudy(path):
    with open(path) as fh:
        toks = fh.read().split()
    nv, ne = int(toks[0]), int(toks[1])
    arr = np.array(toks[2:2 + 3 * ne], dtype=float).reshape(ne, 3)
    return nv, arr[:, 0].astype(np.int32), arr[:, 1].astype(np.int32), arr[:, 2].copy()


def gnp(n, p, weights, seed):
    """G(n, p) with weights drawn from `weights` (array of allowed values;
    an edge drawn with weight 0 is dropped)."""
    rng = np.random.default_rng(seed)
    ei, ej, w = [], [], []
    weights = np.asarray(weights)
    for i in range(1, n + 1):
        for j in range(i + 1, n + 1):
            if rng.random() < p:
                val = weights[rng.integers(len(weights))]
                if val != 0:
                    ei.append(i)
                    ej.append(j)
                    w.append(float(val))
    return n, np.array(ei, np.int32), np.array(ej, np.int32), np.array(w)


def torus2d(r, c, seed):
    """r x c toroidal grid, weights +-1 with equal probability."""
    rng = np.random.default_rng(seed)
    ei, ej = [], []
    idx = lambda a, b: (a % r) * c + (b % c) + 1
    for a in range(r):
        for b in range(c):
            ei.append(idx(a, b)); ej.append(idx(a, b + 1))
            ei.append(idx(a, b)); ej.append(idx(a + 1, b))
    w = rng.choice([-1.0, 1.0], size=len(ei))
    return r * c, np.array(ei, np.int32), np.array(ej, np.int32), w


def torus3d(k, seed):
    """k x k x k toroidal grid, weights round(100 N(0,1))."""
    rng = np.random.default_rng(seed)
    ei, ej = [], []
    idx = lambda a, b, c: ((a % k) * k + (b % k)) * k + (c % k) + 1
    for a in range(k):
        for b in range(k):
            for c in range(k):
                for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    ei.append(idx(a, b, c)); ej.append(idx(a + d[0], b + d[1], c + d[2]))
    w = np.round(100.0 * rng.standard_normal(len(ei)))
    return k ** 3, np.array(ei, np.int32), np.array(ej, np.int32), w


def config2(seed=2):
    """G(100, 0.3), integer weights uniform on {-10..10} without 0."""
    vals = [v for v in range(-10, 11) if v != 0]
    return gnp(100, 0.3, vals, seed)


def config4_graph(seed):
    """G(100, 0.5), weights uniform in {-1, 0, 1}."""
    return gnp(100, 0.5, [-1, 0, 1], seed)


def random_fixings(nfree_total, count, seed, kmin=0, kmax=75):
    """count (xfixed, solX) pairs over nfree_total = N-1 variables with a
    uniformly drawn number of fixed vertices in [kmin, kmax]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        k = int(rng.integers(kmin, kmax + 1))
        xf = np.zeros(nfree_total, np.int32)
        sx = np.zeros(nfree_total, np.int32)
        idx = rng.choice(nfree_total, size=k, replace=False)
        xf[idx] = 1
        sx[idx] = rng.integers(0, 2, size=k)
        out.append((xf, sx))
    return out
