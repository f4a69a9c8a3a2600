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
  mc_1opt                         heuristic.c:100-198  (oracle/_ref/libmadam_ref_heur.so)
  GW_heuristic + update_best      heuristic.c:24-98, 204-222 on the glibc rand() stream
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


# ---- the rounding heuristic: oracle/_ref/libmadam_ref_heur.so is the
# reference's heuristic.c with the evaluateSolution it calls
# (bab_functions.c) and the globals it reads, nothing else

class RefHeur:
    def __init__(self, SPL):
        self.L = C.CDLL(O.REF_HEUR_SO, mode=C.RTLD_LOCAL)
        for name in ("GW_heuristic", "mc_1opt", "evaluateSolution", "runHeuristic"):
            getattr(self.L, name).restype = C.c_double
        self.N = SPL.shape[0]
        self.SPLk = np.asfortranarray(SPL).ravel(order="F").copy()
        self.SP = O.Ref.Problem()
        self.SP.L = self.SPLk.ctypes.data_as(PD)
        self.SP.n = self.N
        C.c_void_p.in_dll(self.L, "SP").value = C.addressof(self.SP)
        C.c_int.in_dll(self.L, "BabPbSize").value = self.N - 1

    def problem(self, Lsub):
        P = O.Ref.Problem()
        keep = np.asfortranarray(Lsub).ravel(order="F").copy()
        P.L = keep.ctypes.data_as(PD)
        P.n = Lsub.shape[0]
        return P, keep

    def set_Z(self, Z):
        self.Zk = np.asfortranarray(Z).ravel(order="F").copy()
        C.c_void_p.in_dll(self.L, "Z").value = self.Zk.ctypes.data
        C.c_int.in_dll(self.L, "M").value = Z.shape[1]


ref_heur = pytest.mark.skipif(not os.path.exists(O.REF_HEUR_SO),
                              reason="oracle/_ref/libmadam_ref_heur.so not built")


def _instance(name):
    from tests.helpers import graph_L
    return graph_L(name)


@ref_heur
@pytest.mark.parametrize("name,nfix", [("g05_60.0", 0), ("g05_60.0", 17), ("w01_100.3", 30)])
def test_mc_1opt_matches_reference(name, nfix):
    """mc_1opt (heuristic.c:100-198): the same local optimum and value from
    the same start, bit for bit, on instances with ties between gains (unit
    weights) and with negative weights."""
    SPL = _instance(name)
    N = SPL.shape[0]
    rng = np.random.default_rng(nfix + 5)
    xf = np.zeros(N - 1, np.int32)
    xf[rng.choice(N - 1, nfix, replace=False)] = 1
    sx = (rng.random(N - 1) < 0.5).astype(np.int32) * xf
    Lsub, _ = O.create_subproblem(SPL, xf, sx)
    R = RefHeur(SPL)
    P, keep = R.problem(Lsub)
    for _ in range(12):
        x0 = rng.choice([-1, 1], size=Lsub.shape[0]).astype(np.int32)
        xr = x0.copy()
        cr = R.L.mc_1opt(xr.ctypes.data_as(C.POINTER(C.c_int)), C.byref(P))
        co, xo = O.one_opt(x0.copy(), Lsub)
        assert cr == co
        assert np.array_equal(xr, xo)


@ref_heur
@pytest.mark.parametrize("name,nfix,seed", [("g05_60.0", 0, 11), ("g05_60.0", 9, 12),
                                            ("w01_100.3", 25, 13)])
def test_gw_heuristic_same_stream_as_reference(name, nfix, seed):
    """GW_heuristic (heuristic.c:24-98) with update_best and
    evaluateSolution: on the same factor Z, node and glibc rand() stream,
    the oracle's GW draws the same hyperplanes and returns the same best
    value and cut, bit for bit."""
    SPL = _instance(name)
    N = SPL.shape[0]
    rng = np.random.default_rng(seed)
    node = O.Node()
    idx = rng.choice(N - 1, nfix, replace=False)
    for i in idx:
        node.xfixed[i] = 1
        node.solX[i] = int(rng.integers(0, 2))
    xf = np.array(node.xfixed[:N - 1], np.int32)
    sx = np.array(node.solX[:N - 1], np.int32)
    Lsub, _ = O.create_subproblem(SPL, xf, sx)
    n = Lsub.shape[0]
    # a factor Z (n x M) of an X like ADMM's: the IPM's primal
    X, _, _, _, _ = O.ipm(Lsub)
    w, V = np.linalg.eigh(X)
    keepc = w > 1e-8
    Z = np.asfortranarray(V[:, keepc] * np.sqrt(w[keepc]))
    R = RefHeur(SPL)
    P, keep = R.problem(Lsub)
    R.set_Z(Z)
    num = N
    xr = np.zeros(N - 1, np.int32)
    libc.srand(seed)
    br = R.L.GW_heuristic(C.byref(R.SP), C.byref(P), C.byref(node),
                          xr.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(num))
    s = O.Solver(SPL, O.default_params())
    s.load_subproblem(xf, sx)
    assert s.s.n == n
    s.mat("Z", n)  # the state's Z holds n x M from its start, like the reference's
    zbuf = np.ctypeslib.as_array(s.s.Z, shape=(n * n,))
    zbuf[:n * Z.shape[1]] = Z.ravel(order="F")
    s.s.M = Z.shape[1]
    xo = np.zeros(N - 1, np.int32)
    libc.srand(seed)
    bo = O.lib().mo_gw(s.st, C.byref(node), xo.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(num))
    assert br == bo
    assert np.array_equal(xr, xo)
    # and the value is the cut's objective (evaluateSolution)
    assert R.L.evaluateSolution(xr.ctypes.data_as(C.POINTER(C.c_int))) == br
