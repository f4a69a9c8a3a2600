"""GPU separation kernels against the oracle.

  triangle separation  (cutting_planes.c:48-127): same list as the oracle,
      bit-exact values, when K is not binding or no ties sit at the K-th
      value; max row value bit-exact
  simulated annealing  (qap_simulated_annealing.c:12-206): fed the same
      random stream, identical cost (bit-exact), k-tuple and draw count
"""
import numpy as np
import pytest

import madam_amd as G
from oracle import oracle_py as O
from tests.helpers import admm_case, graph_L

pytestmark = pytest.mark.gpu


def psd_X(n, seed, rank=5):
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((n, rank))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    X = V @ V.T
    np.fill_diagonal(X, 1.0)
    return np.asfortranarray(X)


@pytest.mark.parametrize("n,seed,K", [(12, 1, 5000), (40, 2, 5000), (40, 2, 37), (97, 3, 1000),
                                      (160, 4, 1600), (225, 5, 2250)])
def test_triangle_separation_matches_oracle(n, seed, K):
    X = psd_X(n, seed)
    mx_g, lst_g = G.tri_separate(X, K, 1e-3)
    mx_o, lst_o = O.tri_separate(X, K, 1e-3)
    assert mx_g == mx_o
    assert len(lst_g) == len(lst_o)
    # the oracle's list order follows its slot history; compare as sets
    key = lambda c: (c[0], c[1], c[2], c[3])
    assert sorted(lst_g, key=key) == sorted(lst_o, key=key)
    assert lst_g == sorted(lst_g, key=key)          # enumeration order


def test_triangle_separation_at_admm_point():
    sol, case = admm_case(graph_L("g05_60.0"))
    X = case["X"]
    for K in (600, 50):
        mx_g, lst_g = G.tri_separate(X, K)
        mx_o, lst_o = O.tri_separate(X, K)
        assert mx_g == mx_o
        key = lambda c: (c[0], c[1], c[2], c[3])
        assert sorted(lst_g, key=key) == sorted(lst_o, key=key)


def _overflow_X(kind, n=60):
    """X whose violated triangle rows crowd the top histogram bin: more than
    the 16384 rows the boundary bin holds (then the exact refinement runs)."""
    rng = np.random.default_rng(8)
    if kind == "tied":          # all C(n,3) type-1 rows equal 0.5 exactly
        X = np.full((n, n), -0.5)
    elif kind == "crowded":     # distinct values within 1e-8 of each other
        X = -0.5 + 1e-9 * rng.standard_normal((n, n))
    else:                       # ties on a few levels plus a higher block
        X = np.full((n, n), -0.5)
        X[:6, :6] = -0.9
    X = (X + X.T) / 2
    np.fill_diagonal(X, 1.0)
    return np.asfortranarray(X)


# (block, 20000) is left out: there the K-th value is tied between rows
# found before and after the list filled, where the reference's slot
# replacement keeps later rows and this library the earlier ones
@pytest.mark.parametrize("kind,K", [("tied", 600), ("tied", 20000), ("crowded", 600),
                                    ("crowded", 20000), ("block", 600)])
def test_triangle_separation_boundary_overflow(kind, K):
    """A boundary bin with more than 16384 rows is selected exactly (value
    descending, then enumeration order), the same set as the oracle's and
    the same list on every call."""
    X = _overflow_X(kind)
    mx_g, lst_g = G.tri_separate(X, K, 1e-3)
    mx_o, lst_o = O.tri_separate(X, K, 1e-3)
    assert mx_g == mx_o
    key = lambda c: (c[0], c[1], c[2], c[3])
    assert len(lst_g) == len(lst_o) == K
    assert sorted(lst_g, key=key) == sorted(lst_o, key=key)
    for _ in range(2):
        assert G.tri_separate(X, K, 1e-3)[1] == lst_g


# n = 15 and 60 read X from its copy in LDS, n = 140 from its lower
# triangle in LDS (an n x n FP64 copy does not fit the 160 KB of LDS of a
# workgroup, half of it does), n = 200 from HBM; "asym": X with its upper
# triangle perturbed, which every layout must ignore (the reference copies
# the lower triangle over the upper first, qap_simulated_annealing.c:14-17)
@pytest.mark.parametrize("n,k,seed,asym", [(15, 5, 11, False), (60, 5, 12, False),
                                           (60, 7, 13, False), (140, 7, 14, False),
                                           (200, 5, 15, False), (60, 5, 16, True),
                                           (140, 5, 17, True), (200, 7, 18, True)])
def test_qap_sa_same_stream_as_oracle(n, k, seed, asym):
    X = psd_X(n, seed)
    if asym:
        noise = np.random.default_rng(seed).standard_normal((n, n))
        X = np.asfortranarray(X + 1e-3 * np.triu(noise, 1))
    ntypes = 3 if k == 5 else 4
    types = [1 + (r % ntypes) for r in range(2 * ntypes)]
    sids = list(range(100, 100 + len(types)))
    cost, perms, draws = G.qap_sa(X, k, types, sids, seed=2020)
    for r, t in enumerate(types):
        H = O.pent_H(t) if k == 5 else O.hept_H(t)
        stream = G.rand31_stream(2020, sids[r], int(draws[r]) + 16)
        c_o, p_o, used = O.qap_sa(H, k, X, stream=stream)
        assert used == draws[r]
        assert c_o == cost[r]
        assert list(p_o) == list(perms[r])
