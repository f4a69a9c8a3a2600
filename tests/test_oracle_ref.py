"""Pins the CPU oracle against the reference's own compiled functions.

oracle/_ref/libmadam_ref.so is built by oracle/Makefile from the reference's
operators.c, cutting_planes.c, qap_simulated_annealing.c, ipm_mc_pk.c, heap.c
and allocate_free.c (the parts of the bounding path that compile with this
image's MKL and MPICH headers; admm.c needs CHOLMOD, which is absent).  Both
libraries run in this process, so glibc rand() is one shared stream: seeding
it with srand() before each side makes the simulated annealing and the
separation lists comparable draw for draw.

Pins established here (bit-exact unless a tolerance is written):
  op_B, op_Bt                     operators.c:62-150, 171-513
  project_SDP (X_+, Q, Z, M)      operators.c:521-619
  triangle separation + update    cutting_planes.c:48-214
  qap simulated annealing         qap_simulated_annealing.c:12-206
  pentagonal / heptagonal lists   cutting_planes.c:224-605
  ipm_mc_pk                       ipm_mc_pk.c:28-347   (1e-10 relative)
Skipped when oracle/_ref is absent (it is built only where the reference tree
exists).
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.skipif(not os.path.exists(O.REF_SO), reason="oracle/_ref not built")

PD = C.POINTER(C.c_double)
libc = C.CDLL(None)


def sym_X(n, seed, scale=0.45):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    X = np.tril(A) + np.tril(A, -1).T
    np.fill_diagonal(X, 1.0)
    return np.asfortranarray(X)


def psd_X(n, seed):
    """A correlation-like PSD matrix (what ADMM produces)."""
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((n, 4))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    X = V @ V.T
    np.fill_diagonal(X, 1.0)
    return np.asfortranarray(X)


def random_cuts(n, rng, ntri=20, npent=10, nhept=8):
    tri = []
    for _ in range(ntri):
        i, j, k = sorted(rng.choice(n, 3, replace=False), reverse=True)
        tri.append((int(rng.integers(1, 5)), int(i), int(j), int(k), float(rng.random())))
    pent = [(int(rng.integers(1, 4)), tuple(int(v) for v in rng.permutation(n)[:5]), float(rng.random()))
            for _ in range(npent)]
    hept = [(int(rng.integers(1, 5)), tuple(int(v) for v in rng.permutation(n)[:7]), float(rng.random()))
            for _ in range(nhept)]
    return tri, pent, hept


@pytest.mark.parametrize("n,seed", [(9, 0), (17, 1), (40, 2)])
def test_op_B_op_Bt_bit_exact(n, seed):
    rng = np.random.default_rng(seed)
    X = sym_X(n, seed)
    tri, pent, hept = random_cuts(n, rng)
    ref = O.Ref(n)
    ref.set_cuts(tri, pent, hept)
    P = ref.problem(n, len(tri), len(pent), len(hept))
    ref.setX(X)
    y_o = O.op_B(X, tri, pent, hept)
    y_r = np.zeros(len(y_o))
    ref.L.op_B(C.byref(P), y_r.ctypes.data_as(PD), ref.bufs["X"].ctypes.data_as(PD))
    assert np.array_equal(y_o, y_r)
    t = rng.standard_normal(len(y_o))
    Bt_o = O.op_Bt(n, t, tri, pent, hept)
    Xr = np.zeros(n * n)
    ref.L.op_Bt(C.byref(P), Xr.ctypes.data_as(PD), t.ctypes.data_as(PD))
    assert np.array_equal(Bt_o, -Xr.reshape(n, n, order="F"))


@pytest.mark.parametrize("n,seed", [(8, 3), (30, 4), (61, 5)])
def test_project_SDP_matches_reference(n, seed):
    W = sym_X(n, seed, scale=0.8)
    ref = O.Ref(n)
    P = ref.problem(n)
    ref.setX(W)
    ref.L.project_SDP(C.byref(P))
    M = C.c_int.in_dll(ref.L, "M").value
    Xr = np.tril(ref.getmat("X", n))
    Qr = np.tril(ref.getmat("Q", n))
    Zr = ref.getmat("Z", n)[:, :M]
    Xp, Q, Z, ev = O.project_psd(W)
    assert Z.shape[1] == M
    assert np.array_equal(np.tril(Xp), Xr)
    assert np.array_equal(np.tril(Q), Qr)
    assert np.array_equal(Z, Zr)
    # the textbook spectral projection agrees to the dsyevr ABSTOL of 1e-8
    w, V = np.linalg.eigh(W)
    Xe = (V * np.maximum(w, 0)) @ V.T
    assert np.abs(Xp - Xe).max() < 1e-7


@pytest.mark.parametrize("n,seed,K", [(12, 6, 5000), (25, 7, 40), (33, 8, 7)])
def test_triangle_separation_and_update(n, seed, K):
    X = sym_X(n, seed, scale=0.6)
    ref = O.Ref(n)
    ref.params.TriIneq = K
    ref.params.violated_TriIneq = 1e-3
    size = C.c_int(0)
    ref.setX(X)
    best_r = ref.L.getViolated_TriangleInequalities(ref.bufs["X"].ctypes.data_as(PD), n,
                                                    ref.tlist, C.byref(size))
    lst_r = [(ref.tlist[q].type, ref.tlist[q].i, ref.tlist[q].j, ref.tlist[q].k, ref.tlist[q].value)
             for q in range(size.value)]
    best_o, lst_o = O.tri_separate(X, K, 1e-3)
    assert best_o == best_r
    assert lst_o == lst_r

    # purge + dedup + append against the reference's update routine
    rng = np.random.default_rng(seed)
    old = [c[:4] + (0.0,) for c in lst_r[: min(len(lst_r), 6)]]
    old += [(1, 4, 2, 1, 0.0), (2, 5, 3, 0, 0.0)]
    duals = rng.choice([0.0, 1e-6, 0.3, 2.0], size=len(old))
    ref.set_cuts(old)
    P = ref.problem(n, len(old))
    nadd, nsub = C.c_int(0), C.c_int(0)
    ref.L.updateTriangleInequalities.restype = C.c_double
    v_r = ref.L.updateTriangleInequalities(C.byref(P), duals.ctypes.data_as(PD), C.byref(nadd),
                                           C.byref(nsub))
    cuts_r = [(ref.cuts[q].type, ref.cuts[q].i, ref.cuts[q].j, ref.cuts[q].k, ref.cuts[q].y)
              for q in range(P.NIneq)]

    SPL = np.zeros((n, n))
    sol = O.Solver(SPL, O.default_params(TriIneq=K, adjust_TriIneq=0))
    sol.s.n = n
    sol.mat("X")[:, :] = X
    sol.set_cuts(old)
    a, r = C.c_int(0), C.c_int(0)
    v_o = O.lib().mo_tri_update(sol.st, duals.ctypes.data_as(PD), C.byref(a), C.byref(r))
    cuts_o = [(sol.s.cuts[q].type, sol.s.cuts[q].i, sol.s.cuts[q].j, sol.s.cuts[q].k, sol.s.cuts[q].y)
              for q in range(sol.s.ntri)]
    assert v_o == v_r and (a.value, r.value) == (nadd.value, nsub.value)
    assert cuts_o == cuts_r


@pytest.mark.parametrize("k,typ,n,seed", [(5, 1, 10, 11), (5, 3, 23, 12), (7, 2, 15, 13), (7, 4, 31, 14)])
def test_qap_simulated_annealing_same_stream(k, typ, n, seed):
    X = psd_X(n, seed)
    ref = O.Ref(n)
    H = O.pent_H(typ) if k == 5 else O.hept_H(typ)
    ref.setX(X)
    perm_r = np.zeros(k, np.int32)
    libc.srand(seed)
    c_r = ref.L.qap_simulated_annealing(H.ctypes.data_as(C.POINTER(C.c_int)), k,
                                        ref.bufs["X"].ctypes.data_as(PD), n,
                                        perm_r.ctypes.data_as(C.POINTER(C.c_int)))
    libc.srand(seed)
    c_o, perm_o, _ = O.qap_sa(H, k, X)
    assert c_o == c_r
    assert np.array_equal(perm_o, perm_r)


def test_H_patterns_match_reference_tables():
    # the reference's static tables (cutting_planes.c:242-244, 438-441)
    ref_pent = [
        [1] * 25,
        [1, -1, -1, -1, -1, -1, 1, 1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1, 1, 1],
        [1, 1, -1, -1, -1, 1, 1, -1, -1, -1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1, -1, -1, 1, 1, 1],
    ]
    for t in range(1, 4):
        assert list(O.pent_H(t)) == ref_pent[t - 1]
    for t in range(1, 5):
        b = np.array([-1 if a < t - 1 else 1 for a in range(7)])
        assert list(O.hept_H(t)) == list(np.outer(b, b).ravel())


@pytest.mark.parametrize("n,seed,trials,K", [(14, 21, 4, 5000), (20, 22, 9, 5)])
def test_pent_hept_separation_lists(n, seed, trials, K):
    X = psd_X(n, seed)
    ref = O.Ref(n)
    ref.params.Pent_Trials = trials
    ref.params.Hepta_Trials = trials
    ref.params.PentIneq = K
    ref.params.HeptaIneq = K
    ref.setX(X)
    for kind in ("pent", "hept"):
        size = C.c_int(0)
        libc.srand(seed)
        if kind == "pent":
            best_r = ref.L.getViolated_PentagonalInequalities(ref.bufs["X"].ctypes.data_as(PD), n,
                                                              ref.plist, C.byref(size))
            lst_r = [(ref.plist[q].type, tuple(ref.plist[q].perm), ref.plist[q].value)
                     for q in range(size.value)]
            arr = (O.Pent * max(1, K))()
            fn = O.lib().mo_pent_separate
        else:
            best_r = ref.L.getViolated_HeptagonalInequalities(ref.bufs["X"].ctypes.data_as(PD), n,
                                                              ref.hlist, C.byref(size))
            lst_r = [(ref.hlist[q].type, tuple(ref.hlist[q].perm), ref.hlist[q].value)
                     for q in range(size.value)]
            arr = (O.Hept * max(1, K))()
            fn = O.lib().mo_hept_separate
        libc.srand(seed)
        sz = C.c_int(0)
        r, _ = O.make_rng(None)
        best_o = fn(X.ctypes.data_as(PD), n, trials, K, arr, C.byref(sz), C.byref(r))
        lst_o = [(arr[q].type, tuple(arr[q].perm), arr[q].value) for q in range(sz.value)]
        assert best_o == best_r
        assert lst_o == lst_r


@pytest.mark.parametrize("path", ["g05_60.0", "w01_100.3"])
def test_ipm_matches_reference(path):
    here = os.path.dirname(__file__)
    nv, ei, ej, w = O.read_rudy(os.path.join(here, "golden", "rudy", path))
    SPL = O.laplacian_from_edges(nv, ei, ej, w)
    L, _ = O.create_subproblem(SPL, np.zeros(nv - 1, np.int32), np.zeros(nv - 1, np.int32))
    n = L.shape[0]
    ref = O.Ref(n)
    X = np.zeros(n * n)
    Z = np.zeros(n * n)
    y = np.zeros(n)
    phi = C.c_double(0)
    Lc = np.asfortranarray(L).ravel(order="F").copy()
    ref.L.ipm_mc_pk(Lc.ctypes.data_as(PD), n, X.ctypes.data_as(PD), y.ctypes.data_as(PD),
                    Z.ctypes.data_as(PD), C.byref(phi), 0)
    Xo, yo, Zo, phio, _ = O.ipm(L)
    assert abs(phio - phi.value) <= 1e-10 * abs(phi.value)
    assert np.abs(yo - y).max() <= 1e-8 * np.abs(y).max()
    assert np.abs(Xo.ravel(order="F") - X).max() <= 1e-8
