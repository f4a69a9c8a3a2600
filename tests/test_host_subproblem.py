"""The library's node reduction (madam_create_subproblem: createSubproblem +
getFixedValue, evaluate.c:39-139), a host function callable without a GPU,
checked on the CPU:

  * against the oracle's restatement: the same PP->L bit for bit and the
    same constant;
  * against the meaning of the reduction: for every completion x of the
    fixings, the Max-Cut objective x^T A x of the whole graph equals
    fixedvalue + z^T (PP->L) z with z = (2 x_free - 1, 1) -- checked over all
    completions of small subproblems.
"""
import itertools

import numpy as np
import pytest

import madam_amd as G
from madam_amd import instances as I
from oracle import oracle_py as O


def fixings(N, nfix, seed):
    rng = np.random.default_rng(seed)
    xf = np.zeros(N - 1, np.int32)
    sx = np.zeros(N - 1, np.int32)
    idx = rng.choice(N - 1, nfix, replace=False)
    xf[idx] = 1
    sx[idx] = rng.integers(0, 2, nfix)
    return xf, sx


@pytest.mark.parametrize("N,nfix,seed", [(8, 0, 1), (12, 4, 2), (14, 10, 3), (40, 17, 4)])
def test_subproblem_matches_oracle_and_the_objective(N, nfix, seed):
    SPL = I.objective_matrix(*I.gnp(N, 0.5, [-3, -1, 2, 5], seed))
    xf, sx = fixings(N, nfix, seed)
    L, fv = G.create_subproblem(SPL, xf, sx)
    Lo, fo = O.create_subproblem(SPL, xf, sx)
    assert np.array_equal(L, Lo)
    assert fv == fo
    n = L.shape[0]
    free = [i for i in range(N - 1) if not xf[i]]
    assert n == len(free) + 1
    A = SPL[:N - 1, :N - 1]
    rng = np.random.default_rng(seed + 7)
    comps = (np.array(list(itertools.product([0, 1], repeat=len(free)))) if len(free) <= 12
             else rng.integers(0, 2, size=(200, len(free))))
    for xv in comps:
        x = sx.astype(float)
        x[free] = xv
        z = np.ones(n)
        z[:len(free)] = 2 * xv - 1
        assert x @ A @ x == fv + z @ L @ z
