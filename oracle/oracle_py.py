"""ctypes bindings for the CPU FP64 oracle (oracle/madam_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ and __graft_entry__.smoke() (the
values the GPU results are compared with) and bench.py's cpu_baseline leg.
The product never imports it.

Also binds oracle/_ref/libmadam_ref.so (the reference's own operators.c,
cutting_planes.c, qap_simulated_annealing.c, ipm_mc_pk.c, heap.c,
allocate_free.c compiled from the reference's sources by oracle/Makefile) when that
library exists; it is used only to pin the oracle.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libmadam_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmadam_ref.so")
REF_HEUR_SO = os.path.join(HERE, "_ref", "libmadam_ref_heur.so")

NMAX = 1024
MAX_CUTS = 50000


class Params(C.Structure):
    """Same layout as Parameters (admm.h:57-89) and mo_params."""
    _fields_ = [
        ("sigma0", C.c_double), ("scaleSigma", C.c_double),
        ("ADMM_itermax", C.c_int), ("tol0", C.c_double),
        ("scaleTol", C.c_double), ("minTol", C.c_double),
        ("triag_iter", C.c_int), ("pent_iter", C.c_int),
        ("hept_iter", C.c_int), ("max_outer_iter", C.c_int),
        ("extra_iter", C.c_int), ("violated_TriIneq", C.c_double),
        ("TriIneq", C.c_int), ("adjust_TriIneq", C.c_int),
        ("PentIneq", C.c_int), ("HeptaIneq", C.c_int),
        ("Pent_Trials", C.c_int), ("Hepta_Trials", C.c_int),
        ("include_Pent", C.c_int), ("include_Hepta", C.c_int),
        ("root", C.c_int), ("use_diff", C.c_int),
        ("branchingStrategy", C.c_int),
    ]


class Tri(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int), ("type", C.c_int),
                ("value", C.c_double), ("y", C.c_double)]


class Pent(C.Structure):
    _fields_ = [("type", C.c_int), ("perm", C.c_int * 5),
                ("value", C.c_double), ("y", C.c_double)]


class Hept(C.Structure):
    _fields_ = [("type", C.c_int), ("perm", C.c_int * 7),
                ("value", C.c_double), ("y", C.c_double)]


class Node(C.Structure):
    """Same layout as BabNode (admm.h:147-155)."""
    _fields_ = [("xfixed", C.c_int * NMAX), ("solX", C.c_int * NMAX),
                ("fracsol", C.c_double * NMAX), ("level", C.c_int),
                ("upper_bound", C.c_double)]


class Rng(C.Structure):
    _fields_ = [("stream", C.POINTER(C.c_int)), ("pos", C.c_long), ("len", C.c_long)]


_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)


class State(C.Structure):
    _fields_ = [
        ("params", Params), ("N", C.c_int), ("SPL", _PD), ("n", C.c_int),
        ("PPL", _PD), ("X", _PD), ("Q", _PD), ("y", _PD), ("Z", _PD), ("W", _PD),
        ("u", _PD), ("t", _PD), ("s", _PD), ("f", C.c_double), ("M", C.c_int),
        ("cuts", C.POINTER(Tri)), ("tri_list", C.POINTER(Tri)),
        ("pcuts", C.POINTER(Pent)), ("pent_list", C.POINTER(Pent)),
        ("hcuts", C.POINTER(Hept)), ("hept_list", C.POINTER(Hept)),
        ("ntri", C.c_int), ("npent", C.c_int), ("nhept", C.c_int),
        ("LB", C.c_double), ("best_x", C.c_int * NMAX), ("diff", C.c_double),
        ("root_bound", C.c_double), ("nodes", C.c_long), ("admm_iters", C.c_long),
        ("rng", Rng), ("lb_override", C.c_double),
    ]


def build() -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _dp(a):
    return a.ctypes.data_as(_PD)


def _ip(a):
    return a.ctypes.data_as(_PI)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.mo_state_new.restype = C.POINTER(State)
        L.mo_state_new.argtypes = [C.c_int, _PD, C.POINTER(Params)]
        L.mo_state_free.argtypes = [C.POINTER(State)]
        for name in ("mo_sdpbound", "mo_evaluate", "mo_bab", "mo_lambda_max",
                     "mo_fixed_value", "mo_tri_separate", "mo_qap_sa",
                     "mo_pent_separate", "mo_hept_separate", "mo_one_opt",
                     "mo_gw_explicit", "mo_tri_update", "mo_pent_update",
                     "mo_hept_update", "mo_evaluate_solution", "mo_gw"):
            getattr(L, name).restype = C.c_double
        L.mo_bab.argtypes = [C.POINTER(State), C.c_long]
        L.mo_pent_H.restype = _PI
        L.mo_hept_H.restype = _PI
        _lib = L
    return _lib


def default_params(**over) -> Params:
    p = Params()
    lib().mo_default_params(C.byref(p))
    for k, v in over.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------- problems

def read_rudy(path):
    """Parse a rudy file: 'nv ne' then 'i j w' lines (process_input.c:197-225)."""
    with open(path) as fh:
        toks = fh.read().split()
    nv, ne = int(toks[0]), int(toks[1])
    arr = np.array(toks[2:2 + 3 * ne], dtype=float).reshape(ne, 3)
    return nv, arr[:, 0].astype(np.int32), arr[:, 1].astype(np.int32), arr[:, 2].copy()


def laplacian_from_edges(nv, ei, ej, w):
    L = np.zeros((nv, nv), order="F")
    ei = np.ascontiguousarray(ei, dtype=np.int32)
    ej = np.ascontiguousarray(ej, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().mo_laplacian_from_edges(C.c_int(nv), C.c_int(len(w)), _ip(ei), _ip(ej), _dp(w),
                                  L.ctypes.data_as(_PD))
    return L


def create_subproblem(SPL, xfixed, solX):
    N = SPL.shape[0]
    SPL = np.asfortranarray(SPL, dtype=np.float64)
    xf = np.zeros(NMAX, np.int32)
    sx = np.zeros(NMAX, np.int32)
    xf[:N - 1] = xfixed
    sx[:N - 1] = solX
    out = np.zeros(N * N)
    n = lib().mo_create_subproblem(C.c_int(N), C.c_int(N - 1), SPL.ctypes.data_as(_PD),
                                   _ip(xf), _ip(sx), _dp(out))
    return np.asfortranarray(out[:n * n].reshape(n, n, order="F")), \
        lib().mo_fixed_value(C.c_int(N), C.c_int(N - 1), SPL.ctypes.data_as(_PD), _ip(xf), _ip(sx))


# ---------------------------------------------------------------- cuts

def tri_array(cuts):
    """cuts: iterable of (type, i, j, k[, y]) -> ctypes array."""
    cuts = list(cuts)
    arr = (Tri * max(1, len(cuts)))()
    for q, c in enumerate(cuts):
        arr[q].type, arr[q].i, arr[q].j, arr[q].k = c[0], c[1], c[2], c[3]
        arr[q].y = c[4] if len(c) > 4 else 0.0
    return arr, len(cuts)


def kgon_array(cls, k, cuts):
    """cuts: iterable of (type, perm tuple[, y])."""
    cuts = list(cuts)
    arr = (cls * max(1, len(cuts)))()
    for q, c in enumerate(cuts):
        arr[q].type = c[0]
        for a in range(k):
            arr[q].perm[a] = c[1][a]
        arr[q].y = c[2] if len(c) > 2 else 0.0
    return arr, len(cuts)


def op_B(X, tri=(), pent=(), hept=()):
    n = X.shape[0]
    X = np.asfortranarray(X, dtype=np.float64)
    ta, nt = tri_array(tri)
    pa, np_ = kgon_array(Pent, 5, pent)
    ha, nh = kgon_array(Hept, 7, hept)
    y = np.zeros(max(1, nt + np_ + nh))
    lib().mo_op_B(C.c_int(n), X.ctypes.data_as(_PD), C.c_int(nt), ta, C.c_int(np_), pa,
                  C.c_int(nh), ha, _dp(y))
    return y[:nt + np_ + nh]


def op_Bt(n, t, tri=(), pent=(), hept=()):
    """Returns B^T(t) as a dense symmetric matrix (op_Bt subtracts it)."""
    X = np.zeros((n, n), order="F")
    ta, nt = tri_array(tri)
    pa, np_ = kgon_array(Pent, 5, pent)
    ha, nh = kgon_array(Hept, 7, hept)
    t = np.ascontiguousarray(t, dtype=np.float64)
    lib().mo_op_Bt(C.c_int(n), X.ctypes.data_as(_PD), C.c_int(nt), ta, C.c_int(np_), pa,
                   C.c_int(nh), ha, _dp(t))
    return -X


def project_psd(W):
    n = W.shape[0]
    W = np.asfortranarray(W, dtype=np.float64)
    Xp = np.zeros((n, n), order="F")
    Q = np.zeros((n, n), order="F")
    Z = np.zeros((n, n), order="F")
    ev = np.zeros(n)
    M = C.c_int(0)
    info = lib().mo_project_psd(C.c_int(n), W.ctypes.data_as(_PD), Xp.ctypes.data_as(_PD),
                                Q.ctypes.data_as(_PD), Z.ctypes.data_as(_PD), _dp(ev),
                                C.byref(M))
    assert info == 0
    return Xp, Q, Z[:, :M.value].copy(), ev[:M.value].copy()


def lambda_max(A):
    A = np.asfortranarray(A, dtype=np.float64)
    return lib().mo_lambda_max(C.c_int(A.shape[0]), A.ctypes.data_as(_PD))


def tri_separate(X, K, thresh=1e-3):
    n = X.shape[0]
    X = np.asfortranarray(X, dtype=np.float64)
    lst = (Tri * max(1, K))()
    size = C.c_int(0)
    best = lib().mo_tri_separate(C.c_int(n), X.ctypes.data_as(_PD), C.c_int(K),
                                 C.c_double(thresh), lst, C.byref(size))
    return best, [(lst[q].type, lst[q].i, lst[q].j, lst[q].k, lst[q].value)
                  for q in range(size.value)]


def make_rng(stream=None):
    r = Rng()
    if stream is None:
        r.stream = None
        r.pos = r.len = 0
        return r, None
    arr = np.ascontiguousarray(stream, dtype=np.int32)
    r.stream = _ip(arr)
    r.pos = 0
    r.len = len(arr)
    return r, arr


def qap_sa(H, k, X, stream=None):
    """One SA run.  stream=None -> glibc rand() (caller seeds via srand)."""
    n = X.shape[0]
    X = np.asfortranarray(X, dtype=np.float64)
    Hh = np.ascontiguousarray(H, dtype=np.int32)
    perm = np.zeros(k, np.int32)
    r, keep = make_rng(stream)
    cost = lib().mo_qap_sa(_ip(Hh), C.c_int(k), X.ctypes.data_as(_PD), C.c_int(n), _ip(perm),
                           C.byref(r))
    return cost, perm, r.pos


def pent_H(t):
    p = lib().mo_pent_H(C.c_int(t))
    return np.array([p[i] for i in range(25)], np.int32)


def hept_H(t):
    p = lib().mo_hept_H(C.c_int(t))
    return np.array([p[i] for i in range(49)], np.int32)


def one_opt(x, L):
    n = L.shape[0]
    x = np.ascontiguousarray(x, dtype=np.int32).copy()
    L = np.asfortranarray(L, dtype=np.float64)
    cost = lib().mo_one_opt(_ip(x), C.c_int(n), L.ctypes.data_as(_PD))
    return cost, x


def gw_explicit(L, Z, planes):
    n = L.shape[0]
    M = Z.shape[1]
    num = planes.shape[0]
    L = np.asfortranarray(L, dtype=np.float64)
    Z = np.asfortranarray(Z, dtype=np.float64)
    planes = np.ascontiguousarray(planes, dtype=np.float64)
    cuts = np.zeros((num, n), np.int32)
    vals = np.zeros(num)
    lib().mo_gw_explicit(C.c_int(n), L.ctypes.data_as(_PD), Z.ctypes.data_as(_PD), C.c_int(M),
                         _dp(planes), C.c_int(num), _ip(cuts), _dp(vals))
    return cuts, vals


def ipm(L):
    n = L.shape[0]
    L = np.asfortranarray(L, dtype=np.float64)
    X = np.zeros((n, n), order="F")
    Z = np.zeros((n, n), order="F")
    y = np.zeros(n)
    phi = C.c_double(0)
    its = C.c_int(0)
    rc = lib().mo_ipm(L.ctypes.data_as(_PD), C.c_int(n), X.ctypes.data_as(_PD), _dp(y),
                      Z.ctypes.data_as(_PD), C.byref(phi), C.byref(its))
    assert rc == 0, rc
    return X, y, Z, phi.value, its.value


class Solver:
    """Whole-instance oracle state (the reference's globals)."""

    def __init__(self, SPL, params=None, stream=None):
        self.SPL = np.asfortranarray(SPL, dtype=np.float64)
        self.N = self.SPL.shape[0]
        p = params if params is not None else default_params()
        self.st = lib().mo_state_new(C.c_int(self.N), self.SPL.ctypes.data_as(_PD), C.byref(p))
        self._stream_keep = None
        if stream is not None:
            r, keep = make_rng(stream)
            self.st.contents.rng = r
            self._stream_keep = keep

    def __del__(self):
        try:
            lib().mo_state_free(self.st)
        except Exception:
            pass

    @property
    def s(self):
        return self.st.contents

    def set_lb(self, lb):
        self.s.LB = lb
        self.s.lb_override = lb

    def bab(self, max_nodes=0):
        return lib().mo_bab(self.st, C.c_long(max_nodes))

    def root_bound(self):
        node = Node()
        return lib().mo_evaluate(self.st, C.byref(node), C.c_int(0)), node

    def evaluate(self, node, rank=1):
        return lib().mo_evaluate(self.st, C.byref(node), C.c_int(rank))

    def load_subproblem(self, xfixed=None, solX=None):
        node = Node()
        if xfixed is not None:
            for i, v in enumerate(xfixed):
                node.xfixed[i] = int(v)
            for i, v in enumerate(solX):
                node.solX[i] = int(v)
        self.s.n = lib().mo_create_subproblem(C.c_int(self.N), C.c_int(self.N - 1),
                                              self.s.SPL, node.xfixed, node.solX, self.s.PPL)
        return node

    def mat(self, name, n=None):
        n = self.s.n if n is None else n
        ptr = getattr(self.s, name)
        return np.ctypeslib.as_array(ptr, shape=(n * n,)).reshape(n, n, order="F")

    def vec(self, name, length):
        ptr = getattr(self.s, name)
        return np.ctypeslib.as_array(ptr, shape=(max(1, length),))[:length]

    def set_cuts(self, tri=(), pent=(), hept=()):
        s = self.s
        for q, c in enumerate(tri):
            s.cuts[q].type, s.cuts[q].i, s.cuts[q].j, s.cuts[q].k = c[0], c[1], c[2], c[3]
            s.cuts[q].y = c[4] if len(c) > 4 else 0.0
        for q, c in enumerate(pent):
            s.pcuts[q].type = c[0]
            for a in range(5):
                s.pcuts[q].perm[a] = c[1][a]
            s.pcuts[q].y = c[2] if len(c) > 2 else 0.0
        for q, c in enumerate(hept):
            s.hcuts[q].type = c[0]
            for a in range(7):
                s.hcuts[q].perm[a] = c[1][a]
            s.hcuts[q].y = c[2] if len(c) > 2 else 0.0
        s.ntri, s.npent, s.nhept = len(tri), len(pent), len(hept)

    def ipm(self):
        n = self.s.n
        phi = C.c_double(0)
        rc = lib().mo_ipm(self.s.PPL, C.c_int(n), self.s.X, self.s.y, self.s.Q, C.byref(phi), None)
        assert rc == 0
        self.s.f = phi.value
        return phi.value

    def admm(self, sigma, tol):
        sg = C.c_double(sigma)
        nbit = C.c_int(0)
        lib().mo_admm(self.st, C.byref(sg), C.c_double(tol), C.byref(nbit))
        return self.s.f, sg.value, nbit.value


# ---------------------------------------------------------------- reference

class Ref:
    """The reference's own compiled functions (oracle/_ref), globals and all."""

    def __init__(self, N):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        self.L = C.CDLL(REF_SO, mode=C.RTLD_LOCAL)
        self.N = N
        L = self.L
        L.getViolated_TriangleInequalities.restype = C.c_double
        L.qap_simulated_annealing.restype = C.c_double
        L.getViolated_PentagonalInequalities.restype = C.c_double
        L.getViolated_HeptagonalInequalities.restype = C.c_double
        self.params = Params.in_dll(L, "params")
        # globals the reference allocates in allocMemory (allocate_free.c:8-46)
        self.bufs = {}
        nn = N * N
        for name, size in (("X", nn), ("Q", nn), ("Z", nn), ("y", N), ("W", N),
                           ("WORK", 26 * N), ("u", 3 * MAX_CUTS), ("t", 3 * MAX_CUTS),
                           ("s", 3 * MAX_CUTS)):
            arr = np.zeros(size)
            self.bufs[name] = arr
            C.c_void_p.in_dll(L, name).value = arr.ctypes.data
        for name, size in (("IWORK", 10 * N), ("ISUPPZ", 2 * N)):
            arr = np.zeros(size, np.int32)
            self.bufs[name] = arr
            C.c_void_p.in_dll(L, name).value = arr.ctypes.data
        C.c_int.in_dll(L, "sizeWORK").value = 26 * N
        C.c_int.in_dll(L, "sizeIWORK").value = 10 * N
        self.cuts = (Tri * MAX_CUTS)()
        self.tlist = (Tri * MAX_CUTS)()
        self.pcuts = (Pent * MAX_CUTS)()
        self.plist = (Pent * MAX_CUTS)()
        self.hcuts = (Hept * MAX_CUTS)()
        self.hlist = (Hept * MAX_CUTS)()
        for name, arr in (("Cuts", self.cuts), ("List", self.tlist), ("Pent_Cuts", self.pcuts),
                          ("Pent_List", self.plist), ("Hepta_Cuts", self.hcuts),
                          ("Hepta_List", self.hlist)):
            C.c_void_p.in_dll(L, name).value = C.addressof(arr)

    class Problem(C.Structure):
        _fields_ = [("L", _PD), ("n", C.c_int), ("NIneq", C.c_int), ("NPentIneq", C.c_int),
                    ("NHeptaIneq", C.c_int)]

    def problem(self, n, ntri=0, npent=0, nhept=0, Lmat=None):
        P = Ref.Problem()
        self._Lkeep = np.zeros(n * n) if Lmat is None else np.asfortranarray(Lmat).ravel(order="F").copy()
        P.L = _dp(self._Lkeep)
        P.n, P.NIneq, P.NPentIneq, P.NHeptaIneq = n, ntri, npent, nhept
        return P

    def setX(self, X):
        n = X.shape[0]
        self.bufs["X"][:n * n] = np.asfortranarray(X).ravel(order="F")

    def getmat(self, name, n):
        return self.bufs[name][:n * n].reshape(n, n, order="F").copy()

    def set_cuts(self, tri=(), pent=(), hept=()):
        for q, c in enumerate(tri):
            self.cuts[q].type, self.cuts[q].i, self.cuts[q].j, self.cuts[q].k = c[0], c[1], c[2], c[3]
            self.cuts[q].y = c[4] if len(c) > 4 else 0.0
        for q, c in enumerate(pent):
            self.pcuts[q].type = c[0]
            for a in range(5):
                self.pcuts[q].perm[a] = c[1][a]
            self.pcuts[q].y = c[2] if len(c) > 2 else 0.0
        for q, c in enumerate(hept):
            self.hcuts[q].type = c[0]
            for a in range(7):
                self.hcuts[q].perm[a] = c[1][a]
            self.hcuts[q].y = c[2] if len(c) > 2 else 0.0
