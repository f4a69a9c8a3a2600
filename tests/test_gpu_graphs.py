"""ADMM check intervals replayed from captured HIP graphs give the same bits
as direct launches.

Each ADMM run below spans many check intervals, so its graph is captured
once and replayed several times, with the host's stop-flag reads and
synchronisations between replays; every result is compared bit for bit
with the same run launched directly.
"""
import numpy as np
import pytest

import madam_amd as G
from tests.helpers import admm_case, graph_L

pytestmark = pytest.mark.gpu


def _run(cases, graphs, **opts):
    G.set_graphs(graphs)
    try:
        return G.ADMM_solver_batch(cases, **opts)
    finally:
        G.set_graphs(False)


def _same(a, b):
    assert a.nbit == b.nbit
    assert a.f == b.f
    assert np.array_equal(a.X, b.X)
    assert np.array_equal(a.y, b.y)


@pytest.mark.parametrize("form", [dict(projection=1, ns_layout=1, cg_rtol=1e-7, cg_adapt=0.01),
                                  dict(projection=1, ns_layout=2, cg_rtol=1e-7, cg_adapt=0.01),
                                  dict(projection=0)])
def test_graph_replay_bit_identical_to_direct_launches(form):
    """(The plain form's block Jacobi solver reads its stop flags on the host
    between sweeps, so it is never captured: the plain case runs every order
    on the one-workgroup Jacobi kernel here.)"""
    if form.get("projection", 0) == 0:
        G.set_bj_min_order(10 ** 6)
    try:
        _graph_replay_case(form)
    finally:
        G.set_bj_min_order(80)


def _graph_replay_case(form):
    cases = []
    for g, kgon in (("g05_60.0", True), ("w01_100.3", True), ("g05_100.0", False)):
        _, c = admm_case(graph_L(g), kgon=kgon)
        c["tol"] = 1e-4
        cases.append(c)
    c0, r0 = G.graph_stats()
    direct = _run(cases, False, **form)
    c1, r1 = G.graph_stats()
    assert (c1, r1) == (c0, r0)
    graphed = _run(cases, True, **form)
    c2, r2 = G.graph_stats()
    print(form, "iterations", [r.nbit for r in direct], "captures", c2 - c1, "replays", r2 - r1)
    assert c2 - c1 >= 1
    assert r2 - r1 >= 3
    for a, b in zip(direct, graphed):
        _same(a, b)


def test_bab_with_graphs_equals_direct():
    """B&B of a G(60, 0.5) instance with weights +-1..+-10 whose tree
    branches (optimum 799): the same optimum over the same nodes with the
    same ADMM iteration count either way."""
    from madam_amd import instances as I
    vals = [v for v in range(-10, 11) if v != 0]
    SPL = I.objective_matrix(*I.gnp(60, 0.5, vals, 6))
    out = []
    for graphs in (False, True):
        G.set_graphs(graphs)
        try:
            B = G.Bounder(SPL, G.default_params(), batch=4, projection=1, cg_rtol=1e-7,
                          cg_adapt=0.01)
            opt = B.bab()
            out.append((opt, int(B.ctx.nodes), int(B.ctx.admm_iters)))
        finally:
            G.set_graphs(False)
    print("direct", out[0], "graphs", out[1], "graph stats", G.graph_stats())
    assert out[0] == out[1]
