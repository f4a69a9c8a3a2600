"""Batched host-device transfers (the pinned staging rings of
madam_engine.h): the same B&B with staging buffers small enough that every
batch wraps the upload ring and splits its downloads, and with the default
64/32 MiB, must take the same path -- same optimum, node count, ADMM and
PCG iteration counts and root bound, bit for bit.

Instances: the branching G(60, 0.5) of the drop-in test (weights +-{1..10},
seed 6; whole tree) and the first 1500 nodes of pw05_150.0 (cut structures
of tens of KB per node, so 1 MiB rings wrap every few dozen nodes and
8 KiB rings send most pieces directly).
"""
import os

import pytest

import madam_amd as G
from madam_amd import instances as I
from tests.helpers import GOLDEN

pytestmark = pytest.mark.gpu

SIZES = [(8192, 8192), (1 << 20, 1 << 20), (1 << 20, 8192)]


def run(SPL, batch, max_nodes, projection):
    kw = dict(projection=1, cg_rtol=1e-7, cg_adapt=0.01, gw_mult=4) if projection else {}
    B = G.Bounder(SPL, G.default_params(), batch=batch, **kw)
    opt = B.bab(max_nodes=max_nodes)
    c = B.ctx
    return (opt, int(c.nodes), int(c.admm_iters), int(c.cg_iters), float(c.root_bound),
            int(c.open_nodes), int(c.admm_calls))


def instance(name):
    if name == "gnp60_s6":
        vals = [v for v in range(-10, 11) if v != 0]
        return I.objective_matrix(*I.gnp(60, 0.5, vals, 6)), 16, 0
    path = os.path.join(GOLDEN, "rudy150", "pw05_150.0")
    return I.objective_matrix(*I.read_rudy(path)), 512, 1500


@pytest.mark.parametrize("name", ["gnp60_s6", "pw05_150.0"])
@pytest.mark.parametrize("projection", [0, 1])
def test_small_staging_rings_take_the_same_path(name, projection):
    SPL, batch, max_nodes = instance(name)
    ref = run(SPL, batch, max_nodes, projection)
    try:
        for up, down in SIZES:
            G.set_staging(up, down)
            got = run(SPL, batch, max_nodes, projection)
            print(name, projection, (up, down), got, "default", ref, flush=True)
            assert got == ref
    finally:
        G.set_staging(0, 0)
