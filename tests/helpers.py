"""Shared test helpers: build ADMM inputs with the oracle."""
import ctypes as C
import os

import numpy as np

from oracle import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
libc = C.CDLL(None)


def graph_L(spec):
    from madam_amd import instances as I
    if isinstance(spec, str):
        nv, ei, ej, w = I.read_rudy(os.path.join(GOLDEN, "rudy", spec))
    else:
        nv, ei, ej, w = spec
    return I.objective_matrix(nv, ei, ej, w)


def admm_case(SPL, kgon=False, seed=7, xfixed=None, solX=None, params=None, hept=True):
    """Oracle state after IPM + first separation: the inputs of the first
    ADMM_solver call of SDPbound (bounding.c:70-128).  kgon: pentagonal
    (and, with hept, heptagonal) cuts separated at the IPM point as well,
    so that the call runs on all cut families; params: overrides of the
    params-file values."""
    p = O.default_params(**(params or {}))
    sol = O.Solver(SPL, p)
    sol.load_subproblem(xfixed, solX)
    sol.ipm()
    a, r = C.c_int(0), C.c_int(0)
    O.lib().mo_tri_update(sol.st, sol.s.s, C.byref(a), C.byref(r))
    if kgon:
        libc.srand(seed)
        O.lib().mo_pent_update(sol.st, sol.s.s, C.c_int(sol.s.ntri), C.byref(a), C.byref(r))
        if hept:
            O.lib().mo_hept_update(sol.st, sol.s.s, C.c_int(sol.s.ntri + sol.s.npent),
                                   C.byref(a), C.byref(r))
    s = sol.s
    tri = [(s.cuts[q].type, s.cuts[q].i, s.cuts[q].j, s.cuts[q].k, s.cuts[q].y) for q in range(s.ntri)]
    pent = [(s.pcuts[q].type, tuple(s.pcuts[q].perm), s.pcuts[q].y) for q in range(s.npent)]
    hept = [(s.hcuts[q].type, tuple(s.hcuts[q].perm), s.hcuts[q].y) for q in range(s.nhept)]
    n = s.n
    case = dict(L=sol.mat("PPL").copy(order="F"), X=sol.mat("X").copy(order="F"),
                Q=sol.mat("Q").copy(order="F"), y=sol.vec("y", n).copy(), tri=tri, pent=pent,
                hept=hept, sigma=p.sigma0, tol=p.tol0, itermax=p.ADMM_itermax,
                scaleSigma=p.scaleSigma)
    return sol, case


def oracle_admm(sol, sigma, tol):
    f, sg, nbit = sol.admm(sigma, tol)
    n = sol.s.n
    m = sol.s.ntri + sol.s.npent + sol.s.nhept
    return dict(f=f, sigma=sg, nbit=nbit, X=sol.mat("X").copy(), y=sol.vec("y", n).copy(),
                s=sol.vec("s", m).copy(), t=sol.vec("t", m).copy())
