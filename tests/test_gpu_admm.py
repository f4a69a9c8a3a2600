"""GPU parity of the ADMM bounding core against the CPU oracle.

Bars (tolerances written here, FP64 throughout):
  op_B, op_Bt                   bit-exact (same order of operations)
  project_SDP                   |X_+ - eigh projection| <= 1e-10 |W|
  lambda_max                    certified: >= numpy eigvalsh, within 1e-9 rel
  ADMM_solver (fixed cut set)   same nbit (+-1 allowed for a stop test that
                                sits on the tolerance), safe bound f within
                                1e-6 relative of the oracle
  batch                         each node equals its single-node run
"""
import itertools

import numpy as np
import pytest

import madam_amd as G
from madam_amd import instances as I
from oracle import oracle_py as O
from tests.helpers import admm_case, graph_L, oracle_admm

pytestmark = pytest.mark.gpu


def sym(n, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    A = np.tril(A) + np.tril(A, -1).T
    return np.asfortranarray(A)


def cuts(n, seed, nt=40, np_=15, nh=10):
    rng = np.random.default_rng(seed)
    tri = []
    for _ in range(nt):
        i, j, k = sorted(rng.choice(n, 3, replace=False), reverse=True)
        tri.append((int(rng.integers(1, 5)), int(i), int(j), int(k), float(rng.random())))
    pent = [(int(rng.integers(1, 4)), tuple(int(v) for v in rng.permutation(n)[:5]), float(rng.random()))
            for _ in range(np_)]
    hept = [(int(rng.integers(1, 5)), tuple(int(v) for v in rng.permutation(n)[:7]), float(rng.random()))
            for _ in range(nh)]
    return tri, pent, hept


@pytest.mark.parametrize("n,seed", [(7, 1), (60, 2), (131, 3)])
def test_op_B_op_Bt_bit_exact_vs_oracle(n, seed):
    X = sym(n, seed)
    tri, pent, hept = cuts(n, seed)
    assert np.array_equal(G.op_B(X, tri, pent, hept), O.op_B(X, tri, pent, hept))
    t = np.random.default_rng(seed).standard_normal(len(tri) + len(pent) + len(hept))
    assert np.array_equal(G.op_Bt(n, t, tri, pent, hept), O.op_Bt(n, t, tri, pent, hept))


@pytest.mark.parametrize("n,seed", [(5, 4), (64, 5), (101, 6), (150, 7), (225, 8)])
def test_project_SDP_vs_eigh(n, seed):
    W = sym(n, seed, 0.7)
    X, Q, Z, M = G.project_SDP(W)
    w, V = np.linalg.eigh(W)
    Xe = (V * np.where(w > 1e-8, w, 0.0)) @ V.T
    scale = np.abs(W).max()
    assert np.abs(X - Xe).max() <= 1e-10 * scale
    assert np.abs(Q - (Xe - W)).max() <= 1e-10 * scale
    assert M == int((w > 1e-8).sum())
    assert np.abs(Z @ Z.T - Xe).max() <= 1e-10 * scale
    # against the oracle's dsyevr path (ABSTOL 1e-8 limits its accuracy)
    Xo, Qo, Zo, evo = O.project_psd(W)
    assert np.abs(X - Xo).max() <= 1e-7


@pytest.mark.parametrize("n,seed", [(3, 9), (100, 10), (200, 11)])
def test_lambda_max_certified(n, seed):
    A = sym(n, seed, 1.0)
    lam = G.lambda_max(A)
    true = np.linalg.eigvalsh(A)[-1]
    assert lam >= true
    assert lam - true <= 1e-9 * max(1.0, abs(true))


@pytest.mark.parametrize("graph,kgon,tol", [("g05_60.0", False, None), ("w01_100.3", True, None),
                                            ("g05_100.0", False, None), ("g05_60.0", False, 1e-3),
                                            ("w01_100.3", True, 1e-3)])
def test_admm_solver_matches_oracle(graph, kgon, tol):
    sol, case = admm_case(graph_L(graph), kgon=kgon)
    if tol is not None:
        case["tol"] = tol
    ref = oracle_admm(sol, case["sigma"], case["tol"])
    got = G.ADMM_solver(**case)
    print(f"{graph}: nbit gpu {got.nbit} oracle {ref['nbit']}; f gpu {got.f:.10f} "
          f"oracle {ref['f']:.10f}; cg {got.cg_iters} sweeps {got.sweeps}")
    assert abs(got.nbit - ref["nbit"]) <= 1
    assert abs(got.f - ref["f"]) <= 1e-6 * abs(ref["f"])
    assert np.abs(got.X - ref["X"]).max() <= 1e-4
    # the safe bound is a valid bound: it dominates the primal value
    assert got.f >= np.sum(case["L"] * got.X) - 1e-6 * abs(got.f)


def test_admm_batch_equals_single():
    cases = []
    for graph in ("g05_60.0", "g05_60.1", "w01_100.3"):
        _, c = admm_case(graph_L(graph))
        cases.append(c)
    batch = G.ADMM_solver_batch(cases)
    for c, b in zip(cases, batch):
        s = G.ADMM_solver(**c)
        assert s.nbit == b.nbit
        assert s.f == b.f
        assert np.array_equal(s.X, b.X)


@pytest.mark.parametrize("n,seed,cluster", [(2, 8, 1), (3, 9, 1), (100, 10, 1), (150, 13, 4),
                                            (190, 14, 1), (200, 11, 6), (512, 12, 1)])
def test_lambda_max_cert_is_an_upper_bound(n, seed, cluster):
    """Estimate + Cholesky certificate (the chosen form's safe bound; the
    estimate by Householder tridiagonalisation for n <= 190 here, by Lanczos
    for 200 and 512): never below lambda_max (eigh), at most 1e-8 |A|_F
    above it, also with a multiple top eigenvalue."""
    A = sym(n, seed, 1.0)
    if cluster > 1:
        w, V = np.linalg.eigh(A)
        w[-cluster:] = w[-1]
        A = np.asfortranarray((V * w) @ V.T)
        A = np.asfortranarray((A + A.T) / 2)
    lam = G.lambda_max_cert(A)
    top = np.linalg.eigvalsh(A)[-1]
    fro = np.linalg.norm(A)
    print(n, "cert", lam, "eigh", top, "excess / |A|_F", (lam - top) / fro)
    assert lam >= top - 1e-13 * fro
    assert lam - top <= 1e-8 * fro


@pytest.mark.parametrize("n", [8, 20])
def test_safe_bound_never_below_the_sdp_value(n):
    """K_n with unit weights: the basic SDP bound is n^2/4 (the Laplacian is
    n I - J, so <L, X> <= n^2 with equality at X = (n I - J) / (n - 1)); one
    triangle cut, inactive at that X, leaves it unchanged.  From the IPM
    start, run to 1e-7, both forms end within 1e-6 of n^2/4, and the chosen
    form's certified bound is never below it (the reference's dsyevr value,
    and so the oracle's, may fall short by up to n x 1e-8)."""
    ei, ej = zip(*itertools.combinations(range(1, n + 1), 2))
    SPL = I.objective_matrix(n, np.array(ei), np.array(ej), np.ones(len(ei)))
    p = O.default_params()
    sol = O.Solver(SPL, p)
    sol.load_subproblem()
    sol.ipm()
    m = sol.s.n
    case = dict(L=sol.mat("PPL").copy(order="F"), X=sol.mat("X").copy(order="F"),
                Q=sol.mat("Q").copy(order="F"), y=sol.vec("y", m).copy(),
                tri=[(1, 2, 1, 0, 0.0)], pent=[], hept=[], sigma=p.sigma0, tol=1e-7,
                itermax=p.ADMM_itermax, scaleSigma=p.scaleSigma)
    want = n * n / 4
    for opts in (dict(projection=1, cg_rtol=1e-7), dict()):
        got = G.ADMM_solver(**case, **opts)
        print(n, opts, "safe bound", got.f, "n^2/4", want, "nbit", got.nbit)
        assert abs(got.f - want) <= 1e-6 * want
        if opts:
            assert got.f >= want
