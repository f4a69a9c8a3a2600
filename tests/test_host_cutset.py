"""The host-side cut structure of libmadam_bound (mh::build_cutset): the CSR
form in which every ADMM kernel reads the cut rows of op_B / op_Bt
(operators.c:62-150, 171-513), checked on the CPU through a small harness
(tests/host/cutset_dump.cpp, built here with hipcc, no device call):

  * shape: cptr counts 3 / 10 / 21 pair entries per triangle / pentagonal /
    heptagonal cut, cscale is 1, 1/2, 1/3;
  * the pair side is the transpose of the cut side, signs included, and
    every active pair is a lower index i + n j with i > j;
  * B(X) summed over the cut side and B^T(t) over the pair side, in the
    streams' order, equal the oracle's op_B / op_Bt bit for bit (the
    oracle's cut maps are pinned bit-exact against the reference's own
    operators.c in tests/test_oracle_ref.py);
  * the 16-bit streams equal the 32-bit ones;
  * a cut with a vertex out of range or repeated is rejected with an error.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc absent")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cutset") / "cutset_dump")
    subprocess.run([HIPCC, "-O1", "-std=c++17", os.path.join(ROOT, "tests", "host", "cutset_dump.cpp"),
                    "-o", exe], check=True, capture_output=True)
    return exe


def run(exe, n, tri, pent, hept):
    lines = [f"{n} {len(tri)} {len(pent)} {len(hept)}"]
    lines += [f"{t} {i} {j} {k}" for t, i, j, k in tri]
    lines += [f"{t} " + " ".join(map(str, p)) for t, p in pent]
    lines += [f"{t} " + " ".join(map(str, p)) for t, p in hept]
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout
    res = {}
    for line in out.splitlines():
        f = line.split()
        if f[0] == "ERR":
            res["err"] = line
        elif f[0] in ("valid", "m", "nact"):
            res[f[0]] = int(f[1])
        else:
            vals = f[2:]
            res[f[0]] = np.array([float(v) for v in vals]) if f[0] == "cscale" else \
                np.array([int(v) for v in vals], dtype=np.int64)
    return res


def random_cuts(n, nt, np_, nh, seed):
    rng = np.random.default_rng(seed)
    tri = []
    for _ in range(nt):
        i, j, k = sorted(rng.choice(n, 3, replace=False), reverse=True)
        tri.append((int(rng.integers(1, 5)), int(i), int(j), int(k)))
    pent = [(int(rng.integers(1, 4)), tuple(int(v) for v in rng.choice(n, 5, replace=False)))
            for _ in range(np_)]
    hept = [(int(rng.integers(1, 5)), tuple(int(v) for v in rng.choice(n, 7, replace=False)))
            for _ in range(nh)]
    return tri, pent, hept


def idx(e):
    return e if e >= 0 else ~e


@pytest.mark.parametrize("n,nt,np_,nh,seed", [(3, 1, 0, 0, 1), (12, 30, 8, 5, 2),
                                              (40, 200, 60, 40, 3)])
def test_cut_structure_reproduces_the_cut_maps(harness, n, nt, np_, nh, seed):
    tri, pent, hept = random_cuts(n, nt, np_, nh, seed)
    r = run(harness, n, tri, pent, hept)
    assert r["valid"] == 1 and "err" not in r
    m = nt + np_ + nh
    cptr, cent, cact = r["cptr"], r["cent"], r["cact"]
    aptr, aent, apair = r["aptr"], r["aent"], r["apair"]
    # shape
    assert r["m"] == m and len(cptr) == m + 1
    sizes = np.diff(cptr)
    assert list(sizes) == [3] * nt + [10] * np_ + [21] * nh
    assert list(r["cscale"]) == [1.0] * nt + [0.5] * np_ + [1.0 / 3.0] * nh
    # pair side = transpose of the cut side, signs included
    nact = r["nact"]
    assert len(aptr) == nact + 1 and len(apair) == nact
    for lo in apair:
        i, j = lo % n, lo // n
        assert i > j
    from_cuts, from_pairs = [], []
    for c in range(m):
        for e in range(cptr[c], cptr[c + 1]):
            a = idx(cact[e])
            assert apair[a] == idx(cent[e])
            assert (cact[e] >= 0) == (cent[e] >= 0)
            from_cuts.append((a, c, cent[e] >= 0))
    for a in range(nact):
        for e in range(aptr[a], aptr[a + 1]):
            from_pairs.append((a, idx(aent[e]), aent[e] >= 0))
    assert sorted(from_cuts) == sorted(from_pairs)
    # 16-bit streams
    assert np.array_equal(r["aent16"], aent) and np.array_equal(r["cact16"], cact)
    # B(X) and B^T(t) in the streams' order against the oracle, bit for bit
    rng = np.random.default_rng(seed + 10)
    A = rng.standard_normal((n, n))
    X = np.asfortranarray(A + A.T)
    Xf = X.ravel(order="F")
    y = np.zeros(m)
    for c in range(m):
        acc = 0.0
        for e in range(cptr[c], cptr[c + 1]):
            v = Xf[idx(cent[e])]
            acc = acc + v if cent[e] >= 0 else acc - v
        y[c] = r["cscale"][c] * acc
    assert np.array_equal(y, O.op_B(X, tri, pent, hept))
    t = rng.standard_normal(m)
    Bt = np.zeros(n * n)
    for a in range(nact):
        acc = 0.0
        for e in range(aptr[a], aptr[a + 1]):
            c = idx(aent[e])
            temp = (0.5 * r["cscale"][c]) * t[c]
            acc = acc + temp if aent[e] >= 0 else acc - temp
        Bt[apair[a]] = acc
    Bo = O.op_Bt(n, t, tri, pent, hept).ravel(order="F")
    assert np.array_equal(Bt[apair], Bo[apair])
    rest = np.setdiff1d(np.arange(n * n), apair)
    lower = [f for f in rest if f % n > f // n]
    assert not np.any(Bo[lower])  # pairs no cut touches stay zero


@pytest.mark.parametrize("tri,pent", [([(1, 3, 1, 0)], []),      # vertex 3 with n = 3
                                      ([(1, 2, 2, 0)], []),      # repeated vertex
                                      ([], [(1, (0, 1, 2, 2, 1))])])
def test_invalid_cuts_are_rejected(harness, tri, pent):
    r = run(harness, 3 if tri else 5, tri, pent, [])
    assert r["valid"] == 0
    assert "err" in r
