"""madam_amd -- Python host mirror of MADAM's SDP bounding interface over the
MI355X C ABI (include/madam_bound.h, libmadam_bound.so).

The functions keep the reference's names and argument meaning:
  op_B, op_Bt, project_SDP           operators.c:62, 171, 521
  ADMM_solver                        admm.c:49
  lambda_max                         the dsyevr call of admm.c:731-747
Everything runs in the HIP library; there is no CPU fallback.  Loading
raises if the library is missing, and every call raises if the library
reports an error (no GPU, HIP failure, bad sizes).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmadam_bound.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "madam_bound.h")

_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)


class MadamError(RuntimeError):
    pass


class Tri(C.Structure):
    """madam_tri == Triangle_Inequality (admm.h:93-100)."""
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int), ("type", C.c_int),
                ("value", C.c_double), ("y", C.c_double)]


class Pent(C.Structure):
    """madam_pent == Pentagonal_Inequality (admm.h:103-108)."""
    _fields_ = [("type", C.c_int), ("perm", C.c_int * 5), ("value", C.c_double),
                ("y", C.c_double)]


class Hept(C.Structure):
    """madam_hept == Heptagonal_Inequality (admm.h:112-117)."""
    _fields_ = [("type", C.c_int), ("perm", C.c_int * 7), ("value", C.c_double),
                ("y", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("cg_rtol", C.c_double), ("cg_maxit", C.c_int),
                ("jacobi_sweeps", C.c_int), ("check_every", C.c_int)]


class AdmmJob(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("L", _PD), ("X", _PD), ("y", _PD), ("Q", _PD),
        ("ntri", C.c_int), ("npent", C.c_int), ("nhept", C.c_int),
        ("tri", C.POINTER(Tri)), ("pent", C.POINTER(Pent)), ("hept", C.POINTER(Hept)),
        ("s", _PD), ("t", _PD), ("u", _PD), ("sigma", _PD), ("tol", C.c_double),
        ("itermax", C.c_int), ("scaleSigma", C.c_double),
        ("nbit", C.c_int), ("f", C.c_double), ("lambda_max", C.c_double),
        ("Z", _PD), ("M", C.c_int), ("cg_iters", C.c_int), ("sweeps", C.c_int),
    ]


def exported_symbols():
    """Function names declared in include/madam_bound.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**\s*(madam_\w+)\s*\(",
                                 txt, re.M)))


_lib = None


def lib():
    """Load libmadam_bound.so; raises MadamError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MadamError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.madam_last_error.restype = C.c_char_p
        L.madam_admm_batch.argtypes = [C.c_int, C.POINTER(AdmmJob), C.POINTER(Opts)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MadamError(f"libmadam_bound error {rc}: {lib().madam_last_error().decode()}")


def open_device(device: int = 0):
    _check(lib().madam_open(C.c_int(device)))


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _tri_arr(tri):
    tri = list(tri)
    arr = (Tri * max(1, len(tri)))()
    for q, c in enumerate(tri):
        arr[q].type, arr[q].i, arr[q].j, arr[q].k = int(c[0]), int(c[1]), int(c[2]), int(c[3])
        arr[q].y = float(c[4]) if len(c) > 4 else 0.0
    return arr, len(tri)


def _kgon_arr(cls, k, cuts):
    cuts = list(cuts)
    arr = (cls * max(1, len(cuts)))()
    for q, c in enumerate(cuts):
        arr[q].type = int(c[0])
        for a in range(k):
            arr[q].perm[a] = int(c[1][a])
        arr[q].y = float(c[2]) if len(c) > 2 else 0.0
    return arr, len(cuts)


def op_B(X, tri=(), pent=(), hept=()):
    """B(X) for the cut lists (operators.c:62-150)."""
    X = _f64(X)
    n = X.shape[0]
    ta, nt = _tri_arr(tri)
    pa, npn = _kgon_arr(Pent, 5, pent)
    ha, nh = _kgon_arr(Hept, 7, hept)
    y = np.zeros(max(1, nt + npn + nh))
    _check(lib().madam_op_B(C.c_int(n), X.ctypes.data_as(_PD), C.c_int(nt), ta, C.c_int(npn), pa,
                            C.c_int(nh), ha, y.ctypes.data_as(_PD)))
    return y[:nt + npn + nh]


def op_Bt(n, t, tri=(), pent=(), hept=()):
    """B^T(t) as a dense symmetric matrix (op_Bt subtracts it, operators.c:171)."""
    X = np.zeros((n, n), order="F")
    ta, nt = _tri_arr(tri)
    pa, npn = _kgon_arr(Pent, 5, pent)
    ha, nh = _kgon_arr(Hept, 7, hept)
    t = np.ascontiguousarray(t, dtype=np.float64)
    _check(lib().madam_op_Bt(C.c_int(n), X.ctypes.data_as(_PD), C.c_int(nt), ta, C.c_int(npn), pa,
                             C.c_int(nh), ha, t.ctypes.data_as(_PD)))
    return -X


def project_SDP(W):
    """X_+ (eigenvalues > 1e-8), Q = X_+ - W, Z (n x M), M (operators.c:521)."""
    W = _f64(W)
    n = W.shape[0]
    X = np.zeros((n, n), order="F")
    Q = np.zeros((n, n), order="F")
    Z = np.zeros((n, n), order="F")
    M = C.c_int(0)
    _check(lib().madam_project_sdp(C.c_int(n), W.ctypes.data_as(_PD), X.ctypes.data_as(_PD),
                                   Q.ctypes.data_as(_PD), Z.ctypes.data_as(_PD), C.byref(M)))
    return X, Q, Z[:, :M.value].copy(), M.value


def lambda_max(A):
    """Certified upper bound on lambda_max(A) (admm.c:731-747)."""
    A = _f64(A)
    lam = C.c_double(0)
    _check(lib().madam_lambda_max(C.c_int(A.shape[0]), A.ctypes.data_as(_PD), C.byref(lam)))
    return lam.value


class AdmmResult(dict):
    __getattr__ = dict.__getitem__


def _make_job(L, X, y, Q, tri, pent, hept, sigma, tol, itermax, scaleSigma, keep):
    L = _f64(L)
    n = L.shape[0]
    X = _f64(X).copy(order="F")
    Q = _f64(Q).copy(order="F")
    y = np.ascontiguousarray(y, dtype=np.float64).copy()
    ta, nt = _tri_arr(tri)
    pa, npn = _kgon_arr(Pent, 5, pent)
    ha, nh = _kgon_arr(Hept, 7, hept)
    m = nt + npn + nh
    s = np.zeros(max(1, m))
    t = np.zeros(max(1, m))
    u = np.zeros(max(1, m))
    Z = np.zeros((n, n), order="F")
    sg = np.array([sigma], dtype=np.float64)
    j = AdmmJob()
    j.n = n
    j.L, j.X, j.y, j.Q = (a.ctypes.data_as(_PD) for a in (L, X, y, Q))
    j.ntri, j.npent, j.nhept = nt, npn, nh
    j.tri = C.cast(ta, C.POINTER(Tri))
    j.pent = C.cast(pa, C.POINTER(Pent))
    j.hept = C.cast(ha, C.POINTER(Hept))
    j.s, j.t, j.u = (a.ctypes.data_as(_PD) for a in (s, t, u))
    j.sigma = sg.ctypes.data_as(_PD)
    j.tol = tol
    j.itermax = itermax
    j.scaleSigma = scaleSigma
    j.Z = Z.ctypes.data_as(_PD)
    keep.append((L, X, Q, y, ta, pa, ha, s, t, u, Z, sg))
    return j, dict(X=X, Q=Q, y=y, s=s, t=t, u=u, Z=Z, sigma=sg, m=m)


def _opts(cg_rtol=1e-12, cg_maxit=1000, jacobi_sweeps=40, check_every=16):
    o = Opts()
    o.cg_rtol, o.cg_maxit, o.jacobi_sweeps, o.check_every = cg_rtol, cg_maxit, jacobi_sweeps, check_every
    return o


def ADMM_solver_batch(problems, **opts):
    """Batched ADMM_solver (admm.c:49) on the GPU.

    problems: list of dicts with keys L, X, y, Q, tri, pent, hept, sigma, tol
    and optionally itermax (2000), scaleSigma (1.001).
    Returns a list of AdmmResult(X, y, Q, s, t, u, sigma, nbit, f, Z, M, ...).
    """
    keep, jobs, outs = [], [], []
    for p in problems:
        j, o = _make_job(p["L"], p["X"], p["y"], p["Q"], p.get("tri", ()), p.get("pent", ()),
                         p.get("hept", ()), p["sigma"], p["tol"], p.get("itermax", 2000),
                         p.get("scaleSigma", 1.001), keep)
        jobs.append(j)
        outs.append(o)
    arr = (AdmmJob * len(jobs))(*jobs)
    o = _opts(**opts)
    _check(lib().madam_admm_batch(C.c_int(len(jobs)), arr, C.byref(o)))
    res = []
    for q, out in enumerate(outs):
        jb = arr[q]
        m = out["m"]
        res.append(AdmmResult(X=out["X"], Q=out["Q"], y=out["y"], s=out["s"][:m], t=out["t"][:m],
                              u=out["u"][:m], sigma=float(out["sigma"][0]), nbit=jb.nbit, f=jb.f,
                              lambda_max=jb.lambda_max, Z=out["Z"][:, :jb.M].copy(), M=jb.M,
                              cg_iters=jb.cg_iters, sweeps=jb.sweeps))
    return res


def ADMM_solver(L, X, y, Q, sigma, tol, tri=(), pent=(), hept=(), itermax=2000,
                scaleSigma=1.001, **opts):
    """ADMM_solver (admm.c:49) for one node."""
    return ADMM_solver_batch([dict(L=L, X=X, y=y, Q=Q, tri=tri, pent=pent, hept=hept,
                                   sigma=sigma, tol=tol, itermax=itermax,
                                   scaleSigma=scaleSigma)], **opts)[0]
