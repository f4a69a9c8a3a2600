"""Drop-in checks: MADAM's own program, unchanged, linked with this
library, under MPI with a master and two workers as the reference runs
(main.c:148-154 sends the root's two children to ranks 1 and 2, so fewer
ranks fail on any tree that branches; the workers share the one GPU).

  oracle/_ref/madam_dropin       main.c, bab_functions.c, heap.c,
      evaluate.c, ... with madam_amd/dropin/madam_sdpbound.c in place of
      bounding.c and admm.c: every SDPbound call -- the root on rank 0, the
      workers' nodes with the diff give-up (bounding.c:111) and the trial
      growth (bounding.c:184-187) -- runs on the GPU.
  oracle/_ref/madam_dropin_admm  the reference's own bounding.c
      (cutting-plane loop, heuristic, separation on the CPU) with
      madam_amd/dropin/madam_admm.c in place of admm.c and ipm_mc_pk.c:
      ADMM_solver (admm.h:212) and ipm_mc_pk (admm.h:269) run on the GPU
      with the reference's prototypes and globals.
  oracle/_ref/madam_dropin_ops   the reference's bounding.c (the
      cutting-plane loop's control flow) with every routine it calls
      through the library: madam_admm.c, madam_operators.c,
      madam_cutting_planes.c, madam_heuristic.c and madam_evaluate.c in
      place of admm.c, ipm_mc_pk.c, operators.c,
      qap_simulated_annealing.c, cutting_planes.c, heuristic.c and
      evaluate.c.
  oracle/_ref/madam_ops_check    tests/dropin_ops_check.c: calls op_B,
      op_Bt, project_SDP, diag, Diag and op_Diag of madam_operators.c, and
      the update*/getViolated_* separation entries of
      madam_cutting_planes.c and mc_1opt/GW_heuristic/update_best of
      madam_heuristic.c, through the reference's prototypes and globals
      (admm.h:237-277, global_var.h); checked against the oracle and the
      library below.

All are built by `make -C oracle all` where the reference tree exists.
Instances: g05_60.0 (rudy, params file values; its root prunes: optimum
536) and G(60, 0.5) with weights uniform on +-{1..10}, seed 6 (its tree
branches: the oracle's own B&B needs 7 nodes).  The proven optimum must
equal the oracle's, in both projection forms.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINS = {name: os.path.join(ROOT, "oracle", "_ref", name)
        for name in ("madam_dropin", "madam_dropin_admm", "madam_dropin_ops")}
OPS_CHECK = os.path.join(ROOT, "oracle", "_ref", "madam_ops_check")
MPIEXEC = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
FIELDS = ["sigma0", "scaleSigma", "ADMM_itermax", "tol0", "scaleTol", "minTol", "triag_iter",
          "pent_iter", "hept_iter", "max_outer_iter", "extra_iter", "violated_TriIneq", "TriIneq",
          "adjust_TriIneq", "PentIneq", "HeptaIneq", "Pent_Trials", "Hepta_Trials",
          "include_Pent", "include_Hepta", "root", "use_diff", "branchingStrategy"]


def write_params(path):
    p = O.default_params()
    with open(path, "w") as f:
        for name in FIELDS:
            f.write(f"{name} = {getattr(p, name)}\n")


def branching_instance():
    from madam_amd import instances as I
    vals = [v for v in range(-10, 11) if v != 0]
    return I.gnp(60, 0.5, vals, 6)


def write_rudy(path, nv, ei, ej, w):
    with open(path, "w") as f:
        f.write(f"{nv} {len(ei)}\n")
        for a, b, c in zip(ei, ej, w):
            f.write(f"{int(a)} {int(b)} {int(c)}\n")


@pytest.fixture(scope="module")
def instances():
    """name -> (writer, proven optimum, oracle B&B nodes); the branching
    instance's optimum comes from the oracle's own B&B."""
    from madam_amd import instances as I
    g = branching_instance()
    C.CDLL(None).srand(2020)
    s = O.Solver(I.objective_matrix(*g), O.default_params())
    opt = s.bab()
    assert s.s.nodes > 1
    return {
        "g05_60.0": (lambda p: shutil.copy(os.path.join(ROOT, "tests", "golden", "rudy",
                                                        "g05_60.0"), p), 536.0, 1),
        "gnp60_s6": (lambda p: write_rudy(p, *g), opt, int(s.s.nodes)),
    }


@pytest.mark.skipif(not os.path.exists(MPIEXEC), reason="mpiexec absent")
@pytest.mark.parametrize("binary", sorted(BINS))
@pytest.mark.parametrize("inst", ["g05_60.0", "gnp60_s6"])
@pytest.mark.parametrize("projection", ["0", "1"])
def test_reference_bab_with_gpu_bounding(tmp_path, instances, binary, inst, projection):
    if not os.path.exists(BINS[binary]):
        pytest.skip(f"oracle/_ref/{binary} not built")
    write, optimum, oracle_nodes = instances[inst]
    write(tmp_path / inst)
    write_params(tmp_path / "params")
    env = dict(os.environ, MADAM_PROJECTION=projection, MKL_NUM_THREADS="1")
    print("running", binary, inst, "projection", projection, flush=True)
    r = subprocess.run([MPIEXEC, "-np", "3", BINS[binary], inst, "params"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    out = (tmp_path / f"{inst}.output").read_text()
    m = re.search(r"Maximum value = (-?\d+)", out)
    assert m and float(m.group(1)) == optimum
    rb = float(re.search(r"Root node bound = (-?[0-9.]+)", out).group(1))
    assert rb >= optimum
    nodes = int(re.search(r"Nodes = (\d+)", out).group(1))
    if oracle_nodes > 1:
        # the tree branched: the workers (rank != 0) bounded nodes
        assert nodes > 1


def _cut_case(n, seed, nt=40, np_=15, nh=10):
    """A full (non-symmetric) X -- the operators read its lower triangle --
    with cuts of every kind (types 1-4, 1-3, 1-4 as the reference's) and t."""
    rng = np.random.default_rng(seed)
    X = np.asfortranarray(rng.standard_normal((n, n)) * 0.7)
    tri = []
    for _ in range(nt):
        i, j, k = sorted(rng.choice(n, 3, replace=False), reverse=True)
        tri.append((int(rng.integers(1, 5)), int(i), int(j), int(k)))
    pent = [(int(rng.integers(1, 4)), tuple(int(v) for v in rng.permutation(n)[:5]))
            for _ in range(np_)]
    hept = [(int(rng.integers(1, 5)), tuple(int(v) for v in rng.permutation(n)[:7]))
            for _ in range(nh)]
    t = rng.standard_normal(nt + np_ + nh)
    return X, tri, pent, hept, t


def _write_case(path, X, tri, pent, hept, t):
    n = X.shape[0]
    with open(path, "wb") as f:
        f.write(np.array([n, len(tri), len(pent), len(hept)], np.int32).tobytes())
        f.write(np.asarray(X, np.float64).tobytes(order="F"))
        for ty, i, j, k in tri:
            f.write(np.array([i, j, k, ty], np.int32).tobytes())
        for ty, perm in pent:
            f.write(np.array([ty, *perm], np.int32).tobytes())
        for ty, perm in hept:
            f.write(np.array([ty, *perm], np.int32).tobytes())
        f.write(np.asarray(t, np.float64).tobytes())


class _Reader:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def take(self, dtype, count):
        a = np.frombuffer(self.b, dtype, count, self.o)
        self.o += a.nbytes
        return a

    def mat(self, n, m=None):
        return self.take(np.float64, n * (n if m is None else m)).reshape(
            (n if m is None else m, n)).T


@pytest.mark.skipif(not os.path.exists(OPS_CHECK), reason="oracle/_ref/madam_ops_check not built")
@pytest.mark.parametrize("n,seed", [(7, 1), (60, 2), (131, 3), (200, 4)])
def test_reference_operator_prototypes(tmp_path, n, seed):
    """op_B / op_Bt / project_SDP / diag / Diag / op_Diag with the
    reference's prototypes and side effects (operators.c:32-619):
      op_B          X's lower triangle mirrored to its upper; y bit-identical
                    to the oracle's B(X)
      op_Bt         X - B^T(t), to 1e-13 (the oracle sums B^T(t) from 0)
      project_SDP   X's lower triangle = X_+ (eigh of the lower triangle, to
                    1e-10 max|X|), its upper triangle untouched; Q = X_out -
                    X_in with a zero upper triangle; M = number of
                    eigenvalues above 1e-8, W = those eigenvalues ascending
                    (1e-10), Z Z^T = X_+ (1e-10)
      diag, Diag, op_Diag  exact"""
    X0, tri, pent, hept, t = _cut_case(n, seed)
    _write_case(tmp_path / "in.bin", X0, tri, pent, hept, t)
    print("madam_ops_check", n, seed, flush=True)
    r = subprocess.run([OPS_CHECK, "ops", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode == 0
    m = len(tri) + len(pent) + len(hept)
    R = _Reader(tmp_path / "out.bin")
    Xs = np.tril(X0) + np.tril(X0, -1).T
    tri5 = [(ty, i, j, k, 0.0) for ty, i, j, k in tri]
    y = R.take(np.float64, m)
    assert np.array_equal(y, O.op_B(Xs, tri5, pent, hept))
    assert np.array_equal(R.mat(n), Xs)
    XmB = R.mat(n)
    assert np.abs(XmB - (X0 - O.op_Bt(n, t, tri5, pent, hept))).max() <= 1e-13

    Xo, Qo = R.mat(n), R.mat(n)
    M = int(R.take(np.int32, 1)[0])
    Wv = R.take(np.float64, M)
    Z = R.mat(n, M) if M else np.zeros((n, 0))
    w, V = np.linalg.eigh(Xs)
    Xp = (V * np.where(w > 1e-8, w, 0.0)) @ V.T
    scale = np.abs(Xs).max()
    low = np.tril_indices(n)
    up = np.triu_indices(n, 1)
    assert np.abs(Xo[low] - Xp[low]).max() <= 1e-10 * scale
    assert np.array_equal(Xo[up], X0[up])
    assert np.abs(Qo[low] - (Xp - Xs)[low]).max() <= 1e-10 * scale
    assert not Qo[up].any()
    assert M == int((w > 1e-8).sum())
    assert np.all(np.diff(Wv) >= 0)
    assert np.abs(Wv - w[w > 1e-8]).max() <= 1e-10 * scale
    assert np.abs(Z @ Z.T - Xp).max() <= 1e-10 * scale

    d = R.take(np.float64, n)
    assert np.array_equal(d, np.diag(X0))
    assert np.array_equal(R.mat(n), np.diag(d))
    assert not R.mat(n).any()
    assert R.o == len(R.b)


def _glibc_rands(seed, count):
    libc = C.CDLL("libc.so.6")
    libc.srand(C.c_uint(seed))
    return [int(libc.rand()) for _ in range(count)]


def _kgon_top_k(cost, perms, types, K):
    """The top-K rules of getViolated_Pentagonal/HeptagonalInequalities
    (cutting_planes.c:247-303) over the runs in order: the first K violated
    runs fill the list, a later run below the tracked value replaces the
    tracked slot, and after a replacement the rescan (which looks for a value
    above BIG_NUMBER) leaves the slot where it is with the value BIG_NUMBER."""
    lst, slot, slotval = [], 0, -1e9
    for r, v in enumerate(cost):
        if not v < 1.0:
            continue
        entry = (int(types[r]), tuple(int(a) for a in perms[r]), float(v))
        if len(lst) < K:
            if not lst or v > slotval:
                slotval, slot = v, len(lst)
            lst.append(entry)
        elif v < slotval:
            lst[slot] = entry
            slotval = 1e9
    return lst


@pytest.mark.skipif(not os.path.exists(OPS_CHECK), reason="oracle/_ref/madam_ops_check not built")
@pytest.mark.parametrize("n,seed", [(40, 7), (70, 8)])
def test_reference_separation_and_rounding_prototypes(tmp_path, n, seed):
    """updateTriangle/Pentagonal/HeptagonalInequalities, mc_1opt and
    GW_heuristic (cutting_planes.c:132-605, heuristic.c:24-223) with the
    reference's prototypes and globals:
      triangles  survivors (dual >= 1e-5) in order, then the oracle's
                 separation rows not among them, in enumeration order; the
                 maximum row value bit-exact
      k-gons     survivors, then the list the reference's top-K rules build
                 from the library's annealing runs (each seed is the first
                 glibc rand() after the driver's srand), duplicates dropped
      mc_1opt    the oracle's 1-opt: same cut and value, bit-exact
      GW         the best of the library's rounding cuts by the test's
                 evaluateSolution, in hyperplane order"""
    from madam_amd import instances as I
    import madam_amd as G
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((n, 5))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    Xs = V @ V.T
    np.fill_diagonal(Xs, 1.0)
    X0 = np.asfortranarray(np.tril(Xs) + np.triu(rng.standard_normal((n, n)), 1))  # upper ignored
    mx_o, lst_o = O.tri_separate(Xs, 200, 1e-3)
    key = lambda c: (c[0], c[1], c[2], c[3])
    sep_sorted = sorted(lst_o, key=key)
    tri = [(c[0], c[1], c[2], c[3]) for c in sep_sorted[:5]]  # will be found again
    for _ in range(25):
        i, j, k = sorted(rng.choice(n, 3, replace=False), reverse=True)
        tri.append((int(rng.integers(1, 5)), int(i), int(j), int(k)))
    pent = [(int(rng.integers(1, 4)), tuple(int(v) for v in rng.permutation(n)[:5])) for _ in range(10)]
    hept = [(int(rng.integers(1, 5)), tuple(int(v) for v in rng.permutation(n)[:7])) for _ in range(6)]
    m = len(tri) + len(pent) + len(hept)
    yy = np.where(rng.random(m) < 0.5, 0.0, rng.random(m) + 1e-3)
    yy[:5] = 1.0
    L = np.asfortranarray(I.objective_matrix(*I.gnp(n, 0.5, [-2, -1, 1, 2], seed)))
    Mz = 5
    Z = np.asfortranarray(rng.standard_normal((n, Mz)))
    x0 = rng.choice([-1, 1], size=n).astype(np.int32)
    K3, K5, K7, T5, T7, num, srand_seed = 200, 30, 30, 12, 10, 40, 1234 + seed
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([n, len(tri), len(pent), len(hept), K3, K5, K7, T5, T7, Mz, num,
                          srand_seed], np.int32).tobytes())
        f.write(np.array([1e-3]).tobytes())
        f.write(X0.tobytes(order="F"))
        for ty, i, j, k in tri:
            f.write(np.array([i, j, k, ty], np.int32).tobytes())
        for ty, perm in pent:
            f.write(np.array([ty, *perm], np.int32).tobytes())
        for ty, perm in hept:
            f.write(np.array([ty, *perm], np.int32).tobytes())
        f.write(yy.tobytes())
        f.write(L.tobytes(order="F"))
        f.write(Z.tobytes(order="F"))
        f.write(x0.tobytes())
    print("madam_ops_check sep", n, seed, flush=True)
    r = subprocess.run([OPS_CHECK, "sep", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode == 0
    R = _Reader(tmp_path / "out.bin")
    s5, s7, s_gw = [_glibc_rands(srand_seed + q, 1)[0] for q in (1, 2, 3)]

    # triangles
    v = R.take(np.float64, 1)[0]
    cnt, added, sub = R.take(np.int32, 3)
    got = []
    for _ in range(cnt):
        e = R.take(np.int32, 4)
        got.append((int(e[3]), int(e[0]), int(e[1]), int(e[2]), float(R.take(np.float64, 1)[0])))
    nt = len(tri)
    surv = [(ty, i, j, k, float(yy[c])) for c, (ty, i, j, k) in enumerate(tri) if yy[c] >= 1e-5]
    have = {s_[:4] for s_ in surv}
    new = [(c[0], c[1], c[2], c[3], 0.0) for c in sep_sorted if key(c) not in have]
    assert v == mx_o
    assert sub == nt - len(surv) and added == len(new)
    assert got == surv + new

    # pentagons and heptagons
    Xsym = np.asfortranarray(Xs)
    for k, items, off, K, T, ntypes, sd in ((5, pent, nt, K5, T5, 3, s5),
                                            (7, hept, nt + len(pent), K7, T7, 4, s7)):
        v = R.take(np.float64, 1)[0]
        cnt, added, sub = R.take(np.int32, 3)
        got = []
        for _ in range(cnt):
            ty = int(R.take(np.int32, 1)[0])
            perm = tuple(int(a) for a in R.take(np.int32, k))
            val, yv = R.take(np.float64, 2)
            got.append((ty, perm, float(yv)))
        types = [1 + q % ntypes for q in range(T * ntypes)]
        cost, perms, _ = G.qap_sa(Xsym, k, types, list(range(len(types))), seed=sd)
        lst = _kgon_top_k(cost, perms, types, K)
        surv = [(ty, perm, float(yy[off + c])) for c, (ty, perm) in enumerate(items)
                if yy[off + c] >= 1e-5]
        have = {(ty, perm) for ty, perm, _ in surv}
        new = [(ty, perm, 0.0) for ty, perm, _ in lst if (ty, perm) not in have]
        print(k, "runs", len(types), "violated", int((cost < 1).sum()), "list", len(lst),
              "added", len(new))
        assert v == cost.min()
        assert sub == len(items) - len(surv) and added == len(new)
        assert got == surv + new

    # mc_1opt
    cost = R.take(np.float64, 1)[0]
    x1 = R.take(np.int32, n)
    c_o, x_o = O.one_opt(x0, L)
    assert cost == c_o
    assert np.array_equal(x1, x_o)

    # GW_heuristic
    best = R.take(np.float64, 1)[0]
    xb = R.take(np.int32, n - 1)
    _, cuts = G.gw_rounding(L, Z, num, seed=s_gw, stream=0)
    e_best, e_x = -1e9, None
    for h in range(num):
        sol = [(int(c) + 1) // 2 for c in cuts[h, :n - 1]]
        val = 0.0
        for i in range(n - 1):
            for j in range(n - 1):
                val += L[j, i] * sol[i] * sol[j]
        if e_best < val:
            e_best, e_x = val, sol
    assert best == e_best
    assert list(xb) == e_x
    assert R.o == len(R.b)
