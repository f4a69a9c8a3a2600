"""Mixed-precision projection (sign iteration on FP32 MFMA + FP64 finish).

  MFMA layouts       FP32 product within 2e-6 |C|, FP64 within 1e-13 |C|
  projection         backward error |X_+ - eigh projection| <= 2e-7 sqrt(n) |W|_F
                     (FP32 rounding accumulated over the 36 products)
                     (the bound moves by ~1e-2 of that, measured with the
                     oracle's ADMM under injected projection error)
  ADMM_solver        safe bound within 1e-6 relative of the FP64 oracle at
                     the same cut set (measured ~6e-8)
  SDPbound           the cutting-plane loop re-selects cuts from X after
                     every ADMM call; ~1e-7 differences in X flip near-tied
                     rows at the top-K boundary, so the trajectories part:
                     the root bound is asserted valid (>= optimum) and
                     within 1e-4 relative of the FP64 oracle (measured
                     1.1e-5 on g05_60.0, the sign-iteration bound tighter)
"""
import os

import numpy as np
import pytest

import madam_amd as G
from madam_amd import instances as I
from oracle import oracle_py as O
from tests.helpers import admm_case, graph_L, oracle_admm

pytestmark = pytest.mark.gpu


def sym(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return np.asfortranarray((A + A.T) / 2)


@pytest.mark.parametrize("n,alpha,beta", [(16, 0.0, 1.0), (48, 0.7, -0.3), (112, 1.5, 0.25)])
def test_mfma_layouts(n, alpha, beta):
    A = sym(n, n)
    B = A @ A + 0.5 * A          # commutes with A: the product is symmetric
    C32, C64 = G.mfma_selftest(A.astype(np.float32), B.astype(np.float32), A, B, alpha, beta)
    A32 = A.astype(np.float32).astype(np.float64)
    B32 = B.astype(np.float32).astype(np.float64)
    ref32 = A32 @ (alpha * np.eye(n) + beta * B32)
    ref64 = A @ (alpha * np.eye(n) + beta * B)
    assert np.abs(C32 - ref32).max() <= 2e-6 * np.abs(ref32).max()
    assert np.abs(C64 - ref64).max() <= 1e-13 * np.abs(ref64).max()


@pytest.mark.parametrize("n,seed", [(20, 1), (100, 2), (225, 3)])
def test_ns_projection_backward_error(n, seed):
    W = sym(n, seed)
    # spread the spectrum down to 1e-5 |W|_F like ADMM's W on the torus
    w, V = np.linalg.eigh(W)
    w = np.sign(w) * np.geomspace(1e-5, 1.0, n)[np.argsort(np.argsort(np.abs(w)))]
    W = np.asfortranarray((V * w) @ V.T)
    X, Q, M = G.project_SDP_ns(W)
    Xe = (V * np.maximum(w, 0)) @ V.T
    fro = np.linalg.norm(W)
    err = np.abs(X - Xe).max() / fro
    print(n, "backward error", err)
    assert err <= 2e-7 * np.sqrt(n)
    assert M == int((w > 0).sum())
    assert np.abs(Q - (X - W)).max() <= 1e-12 * fro


# cut solve: exact (1e-12), or as the chosen form runs it (1e-7, and an
# inexact solve while the ADMM error is large: tolerance 0.01 x the error)
@pytest.mark.parametrize("cg", [(1e-12, 0.0), (1e-7, 0.01)])
@pytest.mark.parametrize("graph,kgon,tol", [("g05_60.0", False, 1e-3), ("w01_100.3", True, 1e-3),
                                            ("w01_100.3", True, 1e-5)])
def test_admm_ns_matches_fp64_oracle(graph, kgon, tol, cg):
    sol, case = admm_case(graph_L(graph), kgon=kgon)
    case["tol"] = tol
    ref = oracle_admm(sol, case["sigma"], case["tol"])
    got = G.ADMM_solver(**case, projection=1, cg_rtol=cg[0], cg_adapt=cg[1])
    rel = abs(got.f - ref["f"]) / abs(ref["f"])
    print(graph, cg, "ns f", got.f, "oracle", ref["f"], "rel", rel, "nbit", got.nbit, ref["nbit"])
    assert rel <= 1e-6
    assert abs(got.nbit - ref["nbit"]) <= 3


def test_sdpbound_root_ns_matches_oracle():
    SPL = graph_L("g05_60.0")
    sol = O.Solver(SPL, O.default_params(include_Pent=0, include_Hepta=0))
    sol.set_lb(536.0)
    bo, _ = sol.root_bound()
    B = G.Bounder(SPL, G.default_params(include_Pent=0, include_Hepta=0), LB=536.0, projection=1)
    bg = B.sdpbound([B.node()], ranks=[0])[0]
    print("ns root", bg, "oracle", bo)
    assert bg >= 536.0
    assert abs(bg - bo) <= 1e-4 * abs(bo)


@pytest.mark.parametrize("cg_rtol", [1e-10, 1e-6])
def test_sdpbound_root_ns_stable_across_cut_solves(cg_rtol):
    """Root SDPbound of config 2 (n=100, tri+pent+hept cuts): with a sign
    schedule that lacks the stability margin, these cut-solve tolerances
    steered the ADMM into a W whose sign iteration diverged (NaN bound).
    The bound must be finite, valid against the plain FP64 form within the
    spread of the cutting-plane loop, and need no Jacobi fallback."""
    SPL = I.objective_matrix(*I.config2(seed=2))
    B = G.Bounder(SPL, G.default_params(), projection=1, cg_rtol=cg_rtol)
    bg = B.sdpbound([B.node()], ranks=[0])[0]
    P = G.Bounder(SPL, G.default_params(), projection=0)
    bp = P.sdpbound([P.node()], ranks=[0])[0]
    print("cg_rtol", cg_rtol, "ns", bg, "plain", bp, "admm", B.ctx.admm_iters,
          "fallbacks", B.ctx.ns_fallbacks)
    assert np.isfinite(bg)
    assert abs(bg - bp) <= 2e-3 * abs(bp)
    assert B.ctx.ns_fallbacks == 0
    assert B.ctx.admm_iters < 2000


def _layout_graph(spec):
    if spec.startswith("gnp"):  # G(n, 0.3), weights +-{1..10}: orders past the LDS-resident kernel
        n = int(spec[3:])
        return I.gnp(n, 0.3, [v for v in range(-10, 11) if v != 0], 11)
    return spec


# orders 60 and 100: one-workgroup kernel with the iterates in LDS; 120, 140
# and 150: one LDS operand buffer that each product's results overwrite;
# 180: one LDS buffer refilled from HBM; 200: iterates in HBM
@pytest.mark.parametrize("graph,kgon", [("g05_60.0", False), ("w01_100.3", True),
                                        ("gnp120", False), ("gnp140", False),
                                        ("gnp150", False), ("gnp180", False),
                                        ("gnp200", False)])
def test_tile_parallel_layout_is_bit_identical(graph, kgon):
    """The tile-parallel launch sequences (16 x 16 tiles, and 2 x 2 blocks of
    tiles) compute the same tiles with the same instruction order as the
    one-workgroup-per-node kernel in each of its storage modes: the ADMM
    results must agree bit for bit."""
    sol, case = admm_case(graph_L(_layout_graph(graph)), kgon=kgon)
    case["tol"] = 1e-3
    a = G.ADMM_solver(**case, projection=1, ns_layout=1)
    b = G.ADMM_solver(**case, projection=1, ns_layout=2)
    c = G.ADMM_solver(**case, projection=1, ns_layout=3)
    print(graph, "layout1", a.f, a.nbit, "layout2", b.f, b.nbit, "layout3", c.f, c.nbit)
    for o in (b, c):
        assert a.nbit == o.nbit
        assert a.f == o.f
        assert np.array_equal(a.X, o.X)


def test_batch_of_mixed_orders_is_bit_identical():
    """One launch holding nodes of order 100, 130 and 150 (the storage mode
    is the largest node's: one staged LDS buffer): each node's ADMM results
    equal the tile-parallel layout's, node alone, bit for bit."""
    cases = []
    for g in ("gnp100", "gnp130", "gnp150"):
        _, c = admm_case(graph_L(_layout_graph(g)))
        c["tol"] = 1e-3
        cases.append(c)
    batch = G.ADMM_solver_batch(cases, projection=1, ns_layout=1)
    for c, b in zip(cases, batch):
        s = G.ADMM_solver(**c, projection=1, ns_layout=2)
        print(c["L"].shape[0], "batch", b.f, b.nbit, "tiles alone", s.f, s.nbit)
        assert s.nbit == b.nbit
        assert s.f == b.f
        assert np.array_equal(s.X, b.X)


def test_heuristic_factor_reproduces_X():
    """In the sign-iteration form the rounding heuristic's Z is a pivoted
    Cholesky factor of the final X: Z Z^T = X up to the part of X that is
    not PSD (the projection's FP32 residue, eigenvalues ~1e-6 below 0),
    which a Cholesky factor cannot carry."""
    sol, case = admm_case(graph_L("w01_100.3"), kgon=True)
    case["tol"] = 1e-3
    got = G.ADMM_solver(**case, projection=1)
    Z, X = got.Z, got.X
    w = np.linalg.eigvalsh(X)
    neg = max(0.0, -w.min())
    err = np.abs(Z @ Z.T - X).max()
    print("Z columns", Z.shape[1], "max |ZZ^T - X|", err, "most negative eigenvalue of X", -neg,
          "max |X|", np.abs(X).max())
    assert Z.shape[1] >= 1
    assert err <= 20 * neg + 1e-8 * np.abs(X).max()


def test_layout_timing_and_tuner(tmp_path):
    """madam_time_projection times each layout, and the autotuner writes a
    table that the library loads and looks up."""
    from madam_amd import tune
    for lay in (1, 2, 3):
        assert G.time_projection(100, 2, lay, 1) > 0.0
    out = str(tmp_path / "t.tsv")
    rows = tune.tune(out, nps=[64, 160], nbs=[1, 4], reps=1)
    assert len(rows) == 4
    try:
        G.load_tuning(out)
        for np_, nb, best, ms in rows:
            assert G.tuned_layout(np_, nb) == best
    finally:
        G.load_tuning(G.TUNE_TABLE if os.path.exists(G.TUNE_TABLE) else None)
