"""GPU parity on the two largest workloads.

  15x15 toroidal grid (weights +-1, seed 3), nodes with 0..75 random
  fixings (n 150..225), triangle + pentagonal cuts: the first ADMM_solver
  call of SDPbound (admm.c:49-771, bounding.c:70-128) at the oracle's cut
  set, IPM warm start and cut duals, in both forms, against the FP64
  oracle; then the 64-node batch of SDPbound (bounding.c:25-250) through
  properties that hold whatever cut sets the loops separate -- every bound
  at least the value of a cut of its subproblem, batch and single-node
  bounds within the cutting-plane loop's spread, no sign-iteration
  projection redone in FP64.
  8x8x8 toroidal grid (n = 512, Gaussian weights x100 rounded): the root's
  first ADMM_solver call on all three cut families, capped at 41
  iterations so that the oracle leg stays short, in both forms; the
  certified lambda_max against numpy's and the bound against a cut value.
"""
import numpy as np
import pytest

import madam_amd as G
from madam_amd import instances as I
from oracle import oracle_py as O
from tests.helpers import admm_case, oracle_admm

pytestmark = pytest.mark.gpu

C3_PARAMS = dict(include_Hepta=0, TriIneq=1500, adjust_TriIneq=0, PentIneq=500, use_diff=0)
C5_PARAMS = dict(TriIneq=3000, adjust_TriIneq=0, PentIneq=1000, HeptaIneq=1000)
PLAIN = dict(projection=0)
CHOSEN = dict(projection=1, cg_rtol=1e-7, cg_adapt=0.01)
CHOSEN_EXACT_CG = dict(projection=1, cg_rtol=1e-10)


def c3_instance():
    SPL = I.objective_matrix(*I.torus2d(15, 15, seed=3))
    fix = I.random_fixings(SPL.shape[0] - 1, 64, seed=33, kmin=0, kmax=75)
    return SPL, fix


def cut_value_lower_bound(L, seeds=8):
    """The best z^T L z over 1-opt local searches (heuristic.c:100-206, the
    oracle's) from random +-1 starts: the value of a cut of the subproblem,
    so at most its maximum cut."""
    rng = np.random.default_rng(1)
    best = -np.inf
    for _ in range(seeds):
        x = rng.choice([-1, 1], size=L.shape[0])
        best = max(best, O.one_opt(x, L)[0])
    return best


def _rel(a, b):
    return abs(a - b) / abs(b)


# nodes of the batch by order: the smallest (n = 152), a middle one and the
# unfixed one (n = 225)
@pytest.mark.parametrize("which", ["min", "mid", "max"])
def test_c3_first_admm_call_matches_oracle(which):
    SPL, fix = c3_instance()
    ns = np.array([SPL.shape[0] - int(xf.sum()) for xf, _ in fix])
    q = {"min": int(np.argmin(ns)), "max": int(np.argmax(ns)),
         "mid": int(np.argsort(ns)[len(ns) // 2])}[which]
    xf, sx = fix[q]
    sol, case = admm_case(SPL, kgon=True, xfixed=xf, solX=sx, params=C3_PARAMS, hept=False)
    assert case["L"].shape[0] == ns[q] and len(case["pent"]) > 0
    ref = oracle_admm(sol, case["sigma"], case["tol"])
    for form in (PLAIN, CHOSEN_EXACT_CG):
        got = G.ADMM_solver(**case, **form)
        assert _rel(got.f, ref["f"]) <= 1e-6, (form, got.f, ref["f"])
        assert abs(got.nbit - ref["nbit"]) <= 1
        assert np.isfinite(got.X).all()
    if which == "max":
        # the bound path's settings (inexact early cut solves) on a call run
        # to the params file's minTol
        sol, case = admm_case(SPL, kgon=True, xfixed=xf, solX=sx, params=C3_PARAMS, hept=False)
        case["tol"] = 1e-3
        ref = oracle_admm(sol, case["sigma"], 1e-3)
        got = G.ADMM_solver(**case, **CHOSEN)
        assert _rel(got.f, ref["f"]) <= 1e-6, (got.f, ref["f"])


def test_c3_batch_bounds_are_valid_and_agree_with_single_nodes():
    SPL, fix = c3_instance()
    B = G.Bounder(SPL, G.default_params(**C3_PARAMS), own_incumbent=1, gw_mult=4, **CHOSEN)
    nodes = [B.node(xf, sx) for xf, sx in fix]
    bounds = B.sdpbound(nodes, ranks=[1] * len(nodes))
    assert B.ctx.ns_fallbacks == 0
    assert np.isfinite(bounds).all()
    for q, (xf, sx) in enumerate(fix):
        PPL, fixed = O.create_subproblem(SPL, xf, sx)
        lb = fixed + cut_value_lower_bound(PPL)
        assert bounds[q] >= lb - 1e-9 * abs(lb), (q, bounds[q], lb)
        assert nodes[q].upper_bound == bounds[q]
    # the same nodes one at a time: the loops separate cuts from iterates
    # that differ in the last bits and draw other random streams, so the
    # bounds agree within the loop's spread, not bit for bit
    for q in (0, 17, 42):
        xf, sx = fix[q]
        B1 = G.Bounder(SPL, G.default_params(**C3_PARAMS), own_incumbent=1, gw_mult=4, **CHOSEN)
        b1 = B1.sdpbound([B1.node(xf, sx)], ranks=[1])[0]
        assert _rel(bounds[q], b1) <= 1e-2, (q, bounds[q], b1)


@pytest.fixture(scope="module")
def c5_case():
    SPL = I.objective_matrix(*I.torus3d(8, seed=5))
    sol, case = admm_case(SPL, kgon=True, params=C5_PARAMS)
    case["itermax"] = 40
    sol.s.params.ADMM_itermax = 40
    ref = oracle_admm(sol, case["sigma"], case["tol"])
    return SPL, case, ref


@pytest.mark.parametrize("form,tol", [("plain", 1e-8), ("chosen", 1e-6)])
def test_c5_root_first_admm_call_matches_oracle(c5_case, form, tol):
    SPL, case, ref = c5_case
    assert case["L"].shape[0] == 512
    assert len(case["tri"]) == 3000 and len(case["pent"]) > 0 and len(case["hept"]) > 0
    got = G.ADMM_solver(**case, **(PLAIN if form == "plain" else CHOSEN_EXACT_CG))
    assert got.nbit == ref["nbit"] == 41
    assert _rel(got.f, ref["f"]) <= tol, (got.f, ref["f"])
    # the safe bound's certified lambda_max (admm.c:724-763) against numpy's
    # on the matrix the call ends with: L - Diag(y) - B^T(s)
    n = case["L"].shape[0]
    A = case["L"] - np.diag(got.y) - O.op_Bt(n, got.s, case["tri"], case["pent"], case["hept"])
    A = (A + A.T) / 2
    lam, nrm = np.linalg.eigvalsh(A)[-1], np.linalg.norm(A)
    assert lam - 1e-12 * nrm <= got.lambda_max <= lam + 1e-9 * nrm
    assert abs(got.f - (n * got.lambda_max + got.y.sum() + got.s.sum())) <= 1e-9 * abs(got.f)
    # a valid bound: at least the value of a cut
    assert got.f >= cut_value_lower_bound(case["L"], seeds=4)
