"""The plain form's many-workgroup block Jacobi eigen-solver (madam_bjac.hip)
against numpy's eigh, the one-workgroup Jacobi kernel and the FP64 oracle.

Bars (FP64 throughout):
  projection         |X_+ - eigh projection| <= 1e-10 max|W|, Z Z^T the same,
                     M = number of eigenvalues above 1e-8 (also on
                     multiple and clustered eigenvalues)
  lambda_max         >= eigvalsh's, within 1e-9 relative
  ADMM_solver        the bars of test_gpu_admm: nbit within 1 and f within
                     1e-6 of the oracle (warm-started projections every
                     iteration)
"""
import numpy as np
import pytest

import madam_amd as G
from tests.helpers import admm_case, graph_L, oracle_admm

pytestmark = pytest.mark.gpu

DEFAULT_MIN = 80


@pytest.fixture
def bj_everywhere():
    G.set_bj_min_order(2)
    yield
    G.set_bj_min_order(DEFAULT_MIN)


def sym(n, seed, scale=0.7):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    return np.asfortranarray(np.tril(A) + np.tril(A, -1).T)


def clustered(n, seed, groups):
    """Symmetric matrix whose spectrum has `groups` multiple eigenvalues
    (half positive, half negative, some zero)."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    vals = np.repeat(np.linspace(-3, 3, groups), -(-n // groups))[:n]
    A = (Q * vals) @ Q.T
    return np.asfortranarray((A + A.T) / 2)


@pytest.mark.parametrize("n,kind", [(2, "rand"), (17, "rand"), (33, "rand"), (100, "rand"),
                                    (150, "clus"), (225, "rand"), (300, "clus"), (512, "rand")])
def test_block_jacobi_projection_vs_eigh(bj_everywhere, n, kind):
    W = sym(n, n) if kind == "rand" else clustered(n, n, 7)
    X, Q, Z, M = G.project_SDP(W)
    w, V = np.linalg.eigh(W)
    Xe = (V * np.where(w > 1e-8, w, 0.0)) @ V.T
    scale = np.abs(W).max()
    err = np.abs(X - Xe).max() / scale
    print(n, kind, "projection error / max|W|", err, "M", M)
    assert err <= 1e-10
    assert np.abs(Q - (Xe - W)).max() <= 1e-10 * scale
    assert M == int((w > 1e-8).sum())
    assert np.abs(Z @ Z.T - Xe).max() <= 1e-10 * scale


@pytest.mark.parametrize("n", [40, 130, 512])
def test_block_jacobi_lambda_max(bj_everywhere, n):
    A = sym(n, 100 + n, 1.0)
    lam = G.lambda_max(A)
    true = np.linalg.eigvalsh(A)[-1]
    assert lam >= true
    assert lam - true <= 1e-9 * max(1.0, abs(true))


@pytest.mark.parametrize("graph,kgon,tol", [("g05_60.0", False, None), ("w01_100.3", True, None),
                                            ("w01_100.3", True, 1e-3)])
def test_admm_with_block_jacobi_matches_oracle(bj_everywhere, graph, kgon, tol):
    sol, case = admm_case(graph_L(graph), kgon=kgon)
    if tol is not None:
        case["tol"] = tol
    ref = oracle_admm(sol, case["sigma"], case["tol"])
    got = G.ADMM_solver(**case)
    print(f"{graph}: nbit gpu {got.nbit} oracle {ref['nbit']}; f gpu {got.f:.10f} "
          f"oracle {ref['f']:.10f}; sweeps {got.sweeps} ({got.sweeps / max(1, got.nbit):.2f} "
          f"per projection)")
    assert abs(got.nbit - ref["nbit"]) <= 1
    assert abs(got.f - ref["f"]) <= 1e-6 * abs(ref["f"])
    assert np.abs(got.X - ref["X"]).max() <= 1e-4


def test_block_jacobi_and_jacobi_agree_on_an_admm_call():
    """The same ADMM call (order 101) with either plain eigen-solver: the same
    iteration count and the same bound to 1e-9."""
    sol, case = admm_case(graph_L("w01_100.3"), kgon=True)
    case["tol"] = 1e-4
    G.set_bj_min_order(10 ** 6)
    try:
        a = G.ADMM_solver(**case)
        G.set_bj_min_order(2)
        b = G.ADMM_solver(**case)
    finally:
        G.set_bj_min_order(DEFAULT_MIN)
    print("Jacobi", a.nbit, a.f, a.sweeps, "block Jacobi", b.nbit, b.f, b.sweeps)
    assert abs(a.nbit - b.nbit) <= 1
    assert abs(a.f - b.f) <= 1e-9 * abs(a.f)
