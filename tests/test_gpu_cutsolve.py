"""GPU cut solve ((BB^T + I) t = rhs by PCG, the CHOLMOD solve of
admm.c:486-551) in each of its kernel classes, chosen per node by the LDS
its vectors need: k_cg1 (cut vectors in registers, m <= 4096), k_cg_x
(one-pass products over the expanded rows), k_cg_big with the cut and pair
vectors in LDS, k_cg_big with only the cut vector in LDS (the pair sums in
HBM), and k_cg (all in HBM).  madam_set_cg_lds_cap caps the LDS the cut
solve plans with and can keep nodes off k_cg_x, so that one problem runs
through every class.

Bars (FP64):
  k_cg_big, pair sums in LDS vs in HBM   bit-identical ADMM results (same
                                         operations in the same order)
  any class vs k_cg_big (HBM pairs)      same nbit (+-1), safe bound f
                                         within 1e-9 relative, X within 1e-6
  every class                            f within 1e-6 relative of the FP64
                                         oracle's ADMM_solver
"""
import numpy as np
import pytest

import madam_amd as G
from tests.helpers import admm_case, graph_L, oracle_admm

pytestmark = pytest.mark.gpu


def pairs_of(case):
    pairs = set()
    for (_, i, j, k, _) in case["tri"]:
        pairs |= {(max(a, b), min(a, b)) for a, b in ((i, j), (i, k), (j, k))}
    for (_, p, _) in case["pent"] + case["hept"]:
        pairs |= {(max(a, b), min(a, b)) for q, a in enumerate(p) for b in p[q + 1:]}
    return len(pairs)


def run_capped(case, cap, one_pass=True, **opts):
    G.set_cg_lds_cap(cap, one_pass)
    try:
        return G.ADMM_solver(**case, **opts)
    finally:
        G.set_cg_lds_cap(0, True)


CASES = {
    # m = 6380 cuts (6000 triangles, 180 pentagonal, 200 heptagonal), 4686
    # active pairs: natively k_cg_x
    "large": ("g05_100.0", dict(TriIneq=6000, adjust_TriIneq=0, PentIneq=400, HeptaIneq=200)),
    # a few hundred cuts: natively k_cg1
    "small": ("g05_60.0", None),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("cg", [(1e-10, 0.0), (1e-7, 0.01)])
def test_cut_solve_classes_agree(name, cg):
    graph, params = CASES[name]
    sol, case = admm_case(graph_L(graph), kgon=True, params=params)
    m = len(case["tri"]) + len(case["pent"]) + len(case["hept"])
    na = pairs_of(case)
    both = (m + na + 64) * 8   # cut and pair vectors in LDS
    cut_only = (m + 64) * 8    # only the cut vector
    opts = dict(projection=1, cg_rtol=cg[0], cg_adapt=cg[1])
    runs = {
        # large: k_cg_x; small: k_cg1
        "native": run_capped(case, 0, **opts),
        # large: k_cg_big, pair sums in LDS; small: k_cg1
        "two-pass": run_capped(case, 0, one_pass=False, **opts),
        # both: k_cg_big, pair sums in HBM
        "hbm-pairs": run_capped(case, both - 8, one_pass=False, **opts),
        # both: k_cg_x
        "one-pass": run_capped(case, both - 8, **opts),
        # both: k_cg
        "all-hbm": run_capped(case, cut_only - 8, **opts),
    }
    ref_run = runs["hbm-pairs"]
    for k, got in runs.items():
        print(f"{name} m {m} nact {na} {k}: f {got.f:.12f} nbit {got.nbit} cg {got.cg_iters}")
    if m > 4096:
        two = runs["two-pass"]
        assert two.nbit == ref_run.nbit
        assert two.f == ref_run.f
        assert np.array_equal(two.X, ref_run.X)
    for k, got in runs.items():
        assert abs(got.nbit - ref_run.nbit) <= 1, k
        assert abs(got.f - ref_run.f) <= 1e-9 * abs(ref_run.f), k
        assert np.abs(got.X - ref_run.X).max() <= 1e-6, k
    if cg[1] == 0.0:
        ref = oracle_admm(sol, case["sigma"], case["tol"])
        for k, got in runs.items():
            assert abs(got.f - ref["f"]) <= 1e-6 * abs(ref["f"]), k
