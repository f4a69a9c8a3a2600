"""Pins of the CPU oracle by mathematics, for the functions the reference's
compiled parts do not cover (admm.c needs CHOLMOD; heuristic.c and
bab_functions.c need the rest of the program, MPI included):

  mc_1opt (heuristic.c:100-206)        the returned value is x^T L x of the
      returned cut; no single flip gains more than the stop threshold
      (flipping x_i changes x^T L x by 4 (L_ii - x_i (L x)_i), and the
      search stops when every such gain is <= 0.001); never below the start
  GW rounding + 1-opt (heuristic.c:24-98)
      every returned cut is such a local optimum, its value x^T L x
  lambda_max of the safe bound (admm.c:727-747)
      the largest eigenvalue from numpy's eigvalsh, to dsyevr's absolute
      tolerance 1e-8 (the reference's ABSTOL), and not the second largest
  ADMM (admm.c:49-771)                 on K_n with unit weights, whose basic
      SDP bound is n^2/4 in closed form, the safe bound from the IPM start
      without cuts is within 1e-6 of n^2/4 at tolerance 1e-7 and below it
      by no more than n x 1e-8 (dsyevr's ABSTOL in the reference: its
      lambda_max may fall short by that much)
  ADMM safe bound, SDPbound and B&B (admm.c, bounding.c, bab_functions.c)
      the root bound is >= and the B&B optimum == the brute-force maximum
      cut, on graphs of 10 to 14 vertices

A dropped diagonal term or a wrong sign in the 1-opt gain, a wrong sign in
its running L x update, a mis-sorted eigenvalue, a safe-bound formula
missing a term: each breaks at least one of these.  Runs on the CPU.
"""
import itertools

import numpy as np
import pytest

from madam_amd import instances as I
from oracle import oracle_py as O


def sym(n, seed, diag=True):
    rng = np.random.default_rng(seed)
    A = rng.integers(-5, 6, size=(n, n)).astype(float)
    A = np.tril(A) + np.tril(A, -1).T
    if not diag:
        np.fill_diagonal(A, 0.0)
    return np.asfortranarray(A)


def flip_gains(L, x):
    return np.diag(L) - x * (L @ x)


@pytest.mark.parametrize("n,seed", [(2, 1), (9, 2), (30, 3), (61, 4)])
def test_one_opt_reaches_a_local_optimum(n, seed):
    L = sym(n, seed)
    rng = np.random.default_rng(seed + 100)
    for _ in range(5):
        x0 = rng.choice([-1, 1], size=n).astype(np.int32)
        cost, x = O.one_opt(x0, L)
        xf = x.astype(float)
        assert set(np.unique(x)) <= {-1, 1}
        assert cost == pytest.approx(xf @ L @ xf, rel=1e-12, abs=1e-9)
        assert cost >= x0 @ L @ x0 - 1e-9
        assert flip_gains(L, xf).max() <= 0.001 + 1e-9


def test_one_opt_takes_the_steepest_flip():
    """From a cut where exactly one flip gains (the others lose), 1-opt makes
    that flip and stops: the result is the start with that vertex flipped."""
    n = 8
    L = np.zeros((n, n))
    for i in range(n - 1):
        L[i, i + 1] = L[i + 1, i] = -1.0  # a path: x^T L x is largest for alternating signs
    x0 = np.array([1, -1, 1, -1, -1, -1, 1, -1], dtype=np.int32)  # vertex 4 out of step
    g = flip_gains(L, x0.astype(float))
    assert (g > 0.001).sum() == 1
    cost, x = O.one_opt(x0, np.asfortranarray(L))
    want = x0.copy()
    want[int(np.argmax(g))] *= -1
    assert list(x) == list(want)
    assert cost == want @ L @ want


@pytest.mark.parametrize("n,M,seed", [(12, 3, 5), (40, 6, 6)])
def test_gw_cuts_are_local_optima(n, M, seed):
    L = sym(n, seed)
    rng = np.random.default_rng(seed)
    Z = np.asfortranarray(rng.standard_normal((n, M)))
    planes = rng.random((16, M)) - 0.5
    cuts, vals = O.gw_explicit(L, Z, planes)
    for c, v in zip(cuts, vals):
        cf = c.astype(float)
        assert v == pytest.approx(cf @ L @ cf, rel=1e-12, abs=1e-9)
        assert flip_gains(L, cf).max() <= 0.001 + 1e-9


@pytest.mark.parametrize("n,seed", [(2, 7), (17, 8), (60, 9)])
def test_lambda_max_is_the_largest_eigenvalue(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    A = np.asfortranarray((A + A.T) / 2)
    got = O.lambda_max(A)
    w = np.linalg.eigvalsh(A)
    assert abs(got - w[-1]) <= 1e-8 + 1e-12 * np.linalg.norm(A)
    assert abs(got - w[-1]) < abs(got - w[-2]) if n > 1 else True


def brute_force(SPL):
    N = SPL.shape[0]
    A = SPL[:N - 1, :N - 1]
    X = np.array(list(itertools.product([0, 1], repeat=N - 1)), dtype=float)
    return np.einsum("ij,jk,ik->i", X, A, X).max()


@pytest.mark.parametrize("n,p,weights,seed", [(10, 0.5, [1], 11), (12, 0.6, [-1, 1], 12),
                                              (14, 0.4, list(range(-10, 11)), 13)])
def test_oracle_bound_and_bab_against_brute_force(n, p, weights, seed):
    SPL = I.objective_matrix(*I.gnp(n, p, weights, seed))
    best = brute_force(SPL)
    root, _ = O.Solver(SPL, O.default_params()).root_bound()
    assert root >= best - 1e-9
    opt = O.Solver(SPL, O.default_params()).bab()
    assert opt == best


@pytest.mark.parametrize("n", [8, 11, 20])
def test_admm_basic_sdp_of_complete_graph(n):
    """K_n, unit weights: the Laplacian is n I - J, so <L, X> <= n^2 for
    every feasible X (<J, X> >= 0) with equality at X = (n I - J) / (n - 1):
    the basic SDP bound is n^2 / 4."""
    ei, ej = zip(*itertools.combinations(range(1, n + 1), 2))
    SPL = I.objective_matrix(n, np.array(ei), np.array(ej), np.ones(len(ei)))
    p = O.default_params()
    sol = O.Solver(SPL, p)
    sol.load_subproblem()
    sol.ipm()
    f, _, nbit = sol.admm(p.sigma0, 1e-7)
    print(n, "safe bound", f, "n^2/4", n * n / 4, "iterations", nbit)
    assert f >= n * n / 4 - n * 1e-8
    assert f <= n * n / 4 * (1 + 1e-6)
