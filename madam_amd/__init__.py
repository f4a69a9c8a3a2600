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
# sign-iteration layout table written by madam_amd/tune.py on an MI355X
TUNE_TABLE = os.path.join(HERE, "ns_tune.tsv")

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
                ("jacobi_sweeps", C.c_int), ("check_every", C.c_int), ("projection", C.c_int),
                ("ns_layout", C.c_int), ("cg_adapt", C.c_double)]


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
        if os.path.exists(TUNE_TABLE):  # the autotuner's layout table (tune.py)
            _check(L.madam_load_tuning(TUNE_TABLE.encode()))
    return _lib


def _check(rc):
    if rc != 0:
        raise MadamError(f"libmadam_bound error {rc}: {lib().madam_last_error().decode()}")


def open_device(device: int = 0):
    _check(lib().madam_open(C.c_int(device)))


def time_projection(n, nb, layout, reps=3):
    """Average milliseconds of one sign-iteration projection of nb synthetic
    nodes of order n in `layout` (1 one workgroup per node, 2 tiles, 3 2 x 2
    tiles)."""
    ms = C.c_double()
    _check(lib().madam_time_projection(C.c_int(n), C.c_int(nb), C.c_int(layout), C.c_int(reps),
                                       C.byref(ms)))
    return ms.value


def load_tuning(path=None):
    """Load a layout table (None: clear it; the built-in crossovers)."""
    _check(lib().madam_load_tuning(None if path is None else str(path).encode()))


def tuned_layout(n, active):
    """The loaded table's layout for order n and `active` nodes (0: none)."""
    return int(lib().madam_tuned_layout(C.c_int(n), C.c_int(active)))


def set_cg_lds_cap(nbytes, one_pass=True):
    """Cap the LDS bytes the cut solve plans with (0: the device's), so that
    small problems run the kernels meant for large cut sets, and with
    one_pass=False keep nodes off the one-pass kernel (testing)."""
    _check(lib().madam_set_cg_lds_cap(C.c_long(nbytes), C.c_int(1 if one_pass else 0)))


def set_staging(up_bytes=0, down_bytes=0):
    """Capacities of the pinned staging buffers of batched host-device
    transfers (0: the defaults, 64 and 32 MiB); small values make them wrap
    and split (testing)."""
    _check(lib().madam_set_staging(C.c_long(up_bytes), C.c_long(down_bytes)))


def set_graphs(on=True):
    """Replay recurring check intervals of the ADMM iterations from a
    captured HIP graph, or (on=False, the library's default) launch every
    kernel directly."""
    _check(lib().madam_set_graphs(C.c_int(1 if on else 0)))


def set_bj_min_order(n):
    """Plain-form eigen-solver: orders below n by the one-workgroup Jacobi
    kernel, orders from n on by the many-workgroup block Jacobi solver."""
    _check(lib().madam_set_bj_min_order(C.c_int(int(n))))


def graph_stats():
    """(graphs captured, graph replays) since the library was loaded."""
    c, r = C.c_long(0), C.c_long(0)
    _check(lib().madam_graph_stats(C.byref(c), C.byref(r)))
    return int(c.value), int(r.value)


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


def project_SDP_ns(W):
    """Mixed-precision project_SDP (sign iteration on MFMA): X_+, Q, M."""
    W = _f64(W)
    n = W.shape[0]
    X = np.zeros((n, n), order="F")
    Q = np.zeros((n, n), order="F")
    M = C.c_int(0)
    _check(lib().madam_project_sdp_ns(C.c_int(n), W.ctypes.data_as(_PD), X.ctypes.data_as(_PD),
                                      Q.ctypes.data_as(_PD), C.byref(M)))
    return X, Q, M.value


def lambda_max_cert(A):
    """Certified upper bound on lambda_max(A): a Householder/Sturm estimate
    proved by a Cholesky factorisation (the chosen form's safe bound)."""
    A = _f64(A)
    lam = C.c_double(0)
    _check(lib().madam_lambda_max_cert(C.c_int(A.shape[0]), A.ctypes.data_as(_PD), C.byref(lam)))
    return lam.value


def ns_schedule():
    """madam_ns_schedule: (steps, 3) float32 coefficients of the sign iteration."""
    k = lib().madam_ns_schedule(0, None)
    if k <= 0:
        raise MadamError(lib().madam_last_error().decode())
    out = np.zeros(3 * k, np.float32)
    lib().madam_ns_schedule(C.c_int(k), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out.reshape(k, 3)


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


def _opts(cg_rtol=1e-12, cg_maxit=1000, jacobi_sweeps=40, check_every=16, projection=0,
          ns_layout=0, cg_adapt=0.0):
    """projection: 0 FP64 eigen (Jacobi), 1 mixed-precision sign iteration;
    ns_layout: 0 automatic (the loaded tuning table, else built-in crossovers),
    1 one workgroup per node, 2 tile-parallel, 3 tile-parallel over 2 x 2 tiles;
    cg_adapt: > 0 loosens the PCG tolerance to cg_adapt x the ADMM error."""
    o = Opts()
    o.cg_rtol, o.cg_maxit, o.jacobi_sweeps, o.check_every = cg_rtol, cg_maxit, jacobi_sweeps, check_every
    o.projection = projection
    o.ns_layout = ns_layout
    o.cg_adapt = cg_adapt
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


def tri_separate(X, K, thresh=1e-3):
    """getViolated_TriangleInequalities (cutting_planes.c:48): (max value,
    [(type, i, j, k, value), ...]) in enumeration order."""
    X = _f64(X)
    n = X.shape[0]
    lst = (Tri * max(1, K))()
    size = C.c_int(0)
    mx = C.c_double(0)
    _check(lib().madam_tri_separate(C.c_int(n), X.ctypes.data_as(_PD), C.c_int(K),
                                    C.c_double(thresh), lst, C.byref(size), C.byref(mx)))
    return mx.value, [(lst[q].type, lst[q].i, lst[q].j, lst[q].k, lst[q].value)
                      for q in range(size.value)]


def qap_sa(X, k, types, stream_ids, seed):
    """Independent qap_simulated_annealing runs (qap_simulated_annealing.c:12)."""
    X = _f64(X)
    n = X.shape[0]
    types = np.ascontiguousarray(types, dtype=np.int32)
    sid = np.ascontiguousarray(stream_ids, dtype=np.int64)
    nr = len(types)
    cost = np.zeros(nr)
    perms = np.zeros((nr, k), np.int32)
    draws = np.zeros(nr, np.int64)
    _check(lib().madam_qap_sa(C.c_int(n), X.ctypes.data_as(_PD), C.c_int(k), C.c_int(nr),
                              types.ctypes.data_as(_PI), sid.ctypes.data_as(C.POINTER(C.c_longlong)),
                              C.c_ulonglong(seed), cost.ctypes.data_as(_PD), perms.ctypes.data_as(_PI),
                              draws.ctypes.data_as(C.POINTER(C.c_longlong))))
    return cost, perms, draws


def rand31_stream(seed, stream_id, count):
    """The 31-bit random stream of SA run `stream_id` (splitmix64 based),
    reproduced on the host: value_t = splitmix64(base + t) >> 33 with
    base = splitmix64(seed ^ splitmix64(stream_id))."""
    def sm(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))
    with np.errstate(over="ignore"):
        base = sm(np.uint64(seed) ^ sm(np.uint64(stream_id)))
        t = np.arange(count, dtype=np.uint64)
        return (sm(base + t) >> np.uint64(33)).astype(np.int64).astype(np.int32)


# ---------------------------------------------------------------- bounding

NMAX = 1024


class Params(C.Structure):
    """madam_params == Parameters (admm.h:57-89)."""
    _fields_ = [
        ("sigma0", C.c_double), ("scaleSigma", C.c_double), ("ADMM_itermax", C.c_int),
        ("tol0", C.c_double), ("scaleTol", C.c_double), ("minTol", C.c_double),
        ("triag_iter", C.c_int), ("pent_iter", C.c_int), ("hept_iter", C.c_int),
        ("max_outer_iter", C.c_int), ("extra_iter", C.c_int), ("violated_TriIneq", C.c_double),
        ("TriIneq", C.c_int), ("adjust_TriIneq", C.c_int), ("PentIneq", C.c_int),
        ("HeptaIneq", C.c_int), ("Pent_Trials", C.c_int), ("Hepta_Trials", C.c_int),
        ("include_Pent", C.c_int), ("include_Hepta", C.c_int), ("root", C.c_int),
        ("use_diff", C.c_int), ("branchingStrategy", C.c_int),
    ]


def default_params(**over):
    """The reference's params file (params:1-23)."""
    p = Params(sigma0=1.0, scaleSigma=1.001, ADMM_itermax=2000, tol0=5e-2, scaleTol=0.6,
               minTol=1e-3, triag_iter=3, pent_iter=3, hept_iter=3, max_outer_iter=12,
               extra_iter=5, violated_TriIneq=1e-3, TriIneq=5000, adjust_TriIneq=1,
               PentIneq=5000, HeptaIneq=5000, Pent_Trials=60, Hepta_Trials=50, include_Pent=1,
               include_Hepta=1, root=0, use_diff=1, branchingStrategy=1)
    for k, v in over.items():
        setattr(p, k, v)
    return p


class Node(C.Structure):
    """madam_node == BabNode (admm.h:147-155)."""
    _fields_ = [("xfixed", C.c_int * NMAX), ("solX", C.c_int * NMAX),
                ("fracsol", C.c_double * NMAX), ("level", C.c_int), ("upper_bound", C.c_double)]


class BabCtx(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("SPL", _PD), ("params", Params), ("LB", C.c_double),
        ("best_x", C.c_int * NMAX), ("diff", C.c_double), ("seed", C.c_ulonglong),
        ("stream_ctr", C.c_longlong), ("projection", C.c_int), ("batch", C.c_int),
        ("nodes", C.c_longlong), ("admm_iters", C.c_longlong), ("admm_calls", C.c_longlong),
        ("ipm_iters", C.c_longlong), ("open_nodes", C.c_longlong), ("root_bound", C.c_double),
        ("seconds_ipm", C.c_double), ("seconds_admm", C.c_double), ("seconds_sep", C.c_double),
        ("seconds_heur", C.c_double), ("cg_rtol", C.c_double), ("cg_iters", C.c_longlong),
        ("ns_fallbacks", C.c_longlong), ("verbose", C.c_int), ("ns_layout", C.c_int),
        ("own_incumbent", C.c_int), ("cg_adapt", C.c_double), ("gw_mult", C.c_int),
        ("last_call", C.c_void_p),
    ]


class CallState(C.Structure):
    """madam_call_state: the inputs of a node's latest ADMM call."""
    _fields_ = [
        ("cap_n", C.c_int), ("cap_tri", C.c_int), ("cap_pent", C.c_int), ("cap_hept", C.c_int),
        ("X", _PD), ("y", _PD), ("Q", _PD), ("tri", C.POINTER(Tri)), ("pent", C.POINTER(Pent)),
        ("hept", C.POINTER(Hept)), ("n", C.c_int), ("ntri", C.c_int), ("npent", C.c_int),
        ("nhept", C.c_int), ("sigma", C.c_double), ("tol", C.c_double), ("calls", C.c_int),
    ]


def create_subproblem(SPL, xfixed, solX):
    """createSubproblem + getFixedValue (evaluate.c:39-139) -> (PP->L, fixedvalue)."""
    SPL = _f64(SPL)
    N = SPL.shape[0]
    xf = np.zeros(N, np.int32)
    sx = np.zeros(N, np.int32)
    xf[:N - 1] = xfixed
    sx[:N - 1] = solX
    out = np.zeros(N * N)
    n = C.c_int(0)
    fv = C.c_double(0)
    _check(lib().madam_create_subproblem(C.c_int(N), SPL.ctypes.data_as(_PD), xf.ctypes.data_as(_PI),
                                         sx.ctypes.data_as(_PI), out.ctypes.data_as(_PD), C.byref(n),
                                         C.byref(fv)))
    k = n.value
    return np.asfortranarray(out[:k * k].reshape(k, k, order="F")), fv.value


def create_subproblem_batch(SPL, fixings):
    """createSubproblem (evaluate.c:39-116) for several nodes on the device:
    fixings = [(xfixed, solX), ...] -> [PP->L, ...]."""
    SPL = _f64(SPL)
    N = SPL.shape[0]
    k = len(fixings)
    xf = np.zeros((k, N - 1), np.int32)
    sx = np.zeros((k, N - 1), np.int32)
    for q, (a, b) in enumerate(fixings):
        xf[q] = a
        sx[q] = b
    out = np.zeros(k * N * N)
    n = np.zeros(k, np.int32)
    _check(lib().madam_create_subproblem_batch(C.c_int(N), SPL.ctypes.data_as(_PD), C.c_int(k),
                                               xf.ctypes.data_as(_PI), sx.ctypes.data_as(_PI),
                                               out.ctypes.data_as(_PD), n.ctypes.data_as(_PI)))
    res = []
    for q in range(k):
        m = int(n[q])
        blk = out[q * N * N: q * N * N + m * m]
        res.append(np.asfortranarray(blk.reshape(m, m, order="F")))
    return res


def read_rudy(path, capN=NMAX):
    """readData (process_input.c:182-312) in the library: SP->L of a rudy file."""
    SPL = np.zeros(capN * capN)
    N = C.c_int(0)
    _check(lib().madam_read_rudy(str(path).encode(), C.c_int(capN), C.byref(N),
                                 SPL.ctypes.data_as(_PD)))
    n = N.value
    return np.asfortranarray(SPL[:n * n].reshape(n, n, order="F"))


def ipm(L):
    """ipm_mc_pk (ipm_mc_pk.c:28): X, y, Z, phi, iterations."""
    L = _f64(L)
    n = L.shape[0]
    X = np.zeros((n, n), order="F")
    Z = np.zeros((n, n), order="F")
    y = np.zeros(n)
    phi = C.c_double(0)
    its = C.c_int(0)
    _check(lib().madam_ipm(C.c_int(n), L.ctypes.data_as(_PD), X.ctypes.data_as(_PD), y.ctypes.data_as(_PD),
                           Z.ctypes.data_as(_PD), C.byref(phi), C.byref(its)))
    return X, y, Z, phi.value, its.value


def gw_planes(seed, stream, num, M):
    """The hyperplanes madam_gw_rounding draws (see include/madam_bound.h)."""
    def sm(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))
    out = np.zeros((num, M))
    with np.errstate(over="ignore"):
        for h in range(num):
            key = sm(np.uint64(seed) ^ sm((np.uint64(stream) << np.uint64(24)) + np.uint64(h)))
            j = np.arange(M, dtype=np.uint64)
            u = (sm(key + j) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
            out[h] = u - 0.5
    return out


def gw_rounding(L, Z, num, seed=2020, stream=0):
    """GW_heuristic + mc_1opt (heuristic.c:24-201) on factor Z: (values, cuts)."""
    L = _f64(L)
    n = L.shape[0]
    Z = _f64(Z)
    M = Z.shape[1] if Z.ndim == 2 else 0
    vals = np.zeros(num)
    cuts = np.zeros((num, n), np.int8)
    _check(lib().madam_gw_rounding(C.c_int(n), L.ctypes.data_as(_PD), Z.ctypes.data_as(_PD), C.c_int(M),
                                   C.c_int(num), C.c_ulonglong(seed), C.c_longlong(stream),
                                   vals.ctypes.data_as(_PD), cuts.ctypes.data_as(C.POINTER(C.c_byte))))
    return vals, cuts


class Bounder:
    """SDPbound / branch and bound over one instance (SP->L) on the GPU."""

    def __init__(self, SPL, params=None, seed=2020, batch=1, LB=0.0, projection=0, cg_rtol=0.0,
                 verbose=0, own_incumbent=0, ns_layout=0, cg_adapt=0.0, gw_mult=1):
        self.SPL = _f64(SPL)
        self.ctx = BabCtx()
        self.ctx.N = self.SPL.shape[0]
        self.ctx.SPL = self.SPL.ctypes.data_as(_PD)
        self.ctx.params = params if params is not None else default_params()
        if self.ctx.params.adjust_TriIneq:
            self.ctx.params.TriIneq = self.ctx.N * 10
        self.ctx.seed = seed
        self.ctx.batch = batch
        self.ctx.LB = LB
        self.ctx.projection = projection
        self.ctx.cg_rtol = cg_rtol
        self.ctx.verbose = verbose
        self.ctx.own_incumbent = own_incumbent
        self.ctx.ns_layout = ns_layout
        self.ctx.cg_adapt = cg_adapt
        self.ctx.gw_mult = gw_mult

    def sdpbound(self, nodes, ranks=None):
        k = len(nodes)
        arr = (C.POINTER(Node) * k)(*[C.pointer(x) for x in nodes])
        rk = np.ascontiguousarray(ranks if ranks is not None else [1] * k, dtype=np.int32)
        out = np.zeros(k)
        _check(lib().madam_sdpbound_batch(C.byref(self.ctx), C.c_int(k), arr, rk.ctypes.data_as(_PI),
                                          out.ctypes.data_as(_PD)))
        return out

    def sdpbound_last_call(self, node, rank=0, cap_cuts=50000):
        """SDPbound of one node that also returns the inputs of its last
        ADMM_solver call (madam_call_state) as an ADMM_solver problem dict:
        the subproblem L, the warm start X, y, Q, sigma, tol and the cut
        lists with the duals the call started from."""
        N = self.ctx.N
        X, Q, y = np.zeros(N * N), np.zeros(N * N), np.zeros(N)
        tri, pent, hept = (Tri * cap_cuts)(), (Pent * cap_cuts)(), (Hept * cap_cuts)()
        st = CallState(cap_n=N, cap_tri=cap_cuts, cap_pent=cap_cuts, cap_hept=cap_cuts)
        st.X, st.y, st.Q = (a.ctypes.data_as(_PD) for a in (X, y, Q))
        st.tri = C.cast(tri, C.POINTER(Tri))
        st.pent = C.cast(pent, C.POINTER(Pent))
        st.hept = C.cast(hept, C.POINTER(Hept))
        self.ctx.last_call = C.cast(C.pointer(st), C.c_void_p)
        try:
            b = self.sdpbound([node], ranks=[rank])[0]
        finally:
            self.ctx.last_call = None
        n = st.n
        L, _ = create_subproblem(self.SPL, list(node.xfixed)[:N - 1], list(node.solX)[:N - 1])
        return b, dict(
            L=L, X=np.asfortranarray(X[:n * n].reshape(n, n, order="F")), y=y[:n].copy(),
            Q=np.asfortranarray(Q[:n * n].reshape(n, n, order="F")),
            tri=[(c.type, c.i, c.j, c.k, c.y) for c in tri[:st.ntri]],
            pent=[(c.type, tuple(c.perm), c.y) for c in pent[:st.npent]],
            hept=[(c.type, tuple(c.perm), c.y) for c in hept[:st.nhept]],
            sigma=st.sigma, tol=st.tol, itermax=self.ctx.params.ADMM_itermax,
            scaleSigma=self.ctx.params.scaleSigma, calls=st.calls)

    def node(self, xfixed=None, solX=None):
        x = Node()
        if xfixed is not None:
            for i, v in enumerate(xfixed):
                x.xfixed[i] = int(v)
            for i, v in enumerate(solX):
                x.solX[i] = int(v)
        return x

    def bab(self, max_nodes=0):
        opt = C.c_double(0)
        _check(lib().madam_bab(C.byref(self.ctx), C.c_longlong(max_nodes), C.byref(opt)))
        return opt.value


def mfma_selftest(A32, B32, A64, B64, alpha, beta):
    """madam_mfma_selftest: (C32, C64) products through the MFMA paths."""
    A32 = np.asfortranarray(A32, dtype=np.float32)
    B32 = np.asfortranarray(B32, dtype=np.float32)
    A64 = _f64(A64)
    B64 = _f64(B64)
    np_, n = A32.shape[0], A64.shape[0]
    C32 = np.zeros((np_, np_), np.float32, order="F")
    C64 = np.zeros((n, n), order="F")
    _check(lib().madam_mfma_selftest(C.c_int(np_), A32.ctypes.data_as(C.POINTER(C.c_float)),
                                     B32.ctypes.data_as(C.POINTER(C.c_float)),
                                     C32.ctypes.data_as(C.POINTER(C.c_float)), C.c_int(n),
                                     A64.ctypes.data_as(_PD), B64.ctypes.data_as(_PD),
                                     C64.ctypes.data_as(_PD), C.c_double(alpha), C.c_double(beta)))
    return C32, C64


# ------------------------------------------------------------------ learned warm start

NFEAT = 10


class Mlp(C.Structure):
    _fields_ = [("nin", C.c_int), ("nh", C.c_int), ("params", C.POINTER(C.c_float))]


_PF = C.POINTER(C.c_float)


def mlp_nparams(nin=NFEAT):
    return lib().madam_mlp_nparams(C.c_int(nin))


def mlp_init(nin=NFEAT, seed=0):
    """FP32 parameters: Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    H = 64
    parts = [rng.uniform(-1, 1, nin * H) * np.sqrt(6.0 / (nin + H)), np.zeros(H),
             rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / (2 * H)), np.zeros(H),
             rng.uniform(-1, 1, H) * np.sqrt(6.0 / (H + 1)), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == mlp_nparams(nin)
    return p


def _net(params, nin):
    params = np.ascontiguousarray(params, dtype=np.float32)
    net = Mlp()
    net.nin, net.nh, net.params = nin, 64, params.ctypes.data_as(_PF)
    return net, params


def mlp_forward(params, feat):
    feat = np.ascontiguousarray(feat, dtype=np.float32)
    net, keep = _net(params, feat.shape[1])
    out = np.zeros(feat.shape[0], np.float32)
    _check(lib().madam_mlp_forward(C.byref(net), C.c_int(feat.shape[0]), feat.ctypes.data_as(_PF),
                                   out.ctypes.data_as(_PF)))
    return out


def mlp_grad(params, feat, target):
    feat = np.ascontiguousarray(feat, dtype=np.float32)
    target = np.ascontiguousarray(target, dtype=np.float32)
    net, keep = _net(params, feat.shape[1])
    g = np.zeros(keep.size, np.float32)
    loss = C.c_float(0)
    _check(lib().madam_mlp_grad(C.byref(net), C.c_int(feat.shape[0]), feat.ctypes.data_as(_PF),
                                target.ctypes.data_as(_PF), g.ctypes.data_as(_PF), C.byref(loss)))
    return g, loss.value


def mlp_train(params, feat, target, epochs=1, batch=4096, lr=1e-3, seed=0):
    """Adam on the GPU; returns (trained parameters, per-step losses)."""
    feat = np.ascontiguousarray(feat, dtype=np.float32)
    target = np.ascontiguousarray(target, dtype=np.float32)
    net, p = _net(np.array(params, dtype=np.float32, copy=True), feat.shape[1])
    N = feat.shape[0]
    steps = (N // min(batch, N)) * epochs
    losses = np.zeros(steps, np.float32)
    _check(lib().madam_mlp_train(C.byref(net), C.c_longlong(N), feat.ctypes.data_as(_PF),
                                 target.ctypes.data_as(_PF), C.c_int(epochs), C.c_int(batch),
                                 C.c_float(lr), C.c_ulonglong(seed), losses.ctypes.data_as(_PF)))
    return p, losses


def pair_features(L, mu=None, sd=None):
    L = _f64(L)
    n = L.shape[0]
    out = np.zeros((n * (n - 1) // 2, NFEAT), np.float32)
    mu_p = sd_p = None
    if mu is not None and sd is not None:
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        sd = np.ascontiguousarray(sd, dtype=np.float32)
        mu_p, sd_p = mu.ctypes.data_as(_PF), sd.ctypes.data_as(_PF)
    _check(lib().madam_pair_features(C.c_int(n), L.ctypes.data_as(_PD), mu_p, sd_p,
                                     out.ctypes.data_as(_PF)))
    return out


def warm_start(L, pred):
    """(X, y, Q) for ADMM_solver from predicted pair values (madam_warm_start)."""
    L = _f64(L)
    n = L.shape[0]
    pred = np.ascontiguousarray(pred, dtype=np.float32)
    X = np.zeros((n, n), order="F")
    Q = np.zeros((n, n), order="F")
    y = np.zeros(n)
    _check(lib().madam_warm_start(C.c_int(n), L.ctypes.data_as(_PD), pred.ctypes.data_as(_PF),
                                  X.ctypes.data_as(_PD), y.ctypes.data_as(_PD),
                                  Q.ctypes.data_as(_PD)))
    return X, y, Q
