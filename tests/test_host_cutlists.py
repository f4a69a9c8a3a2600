"""The cut-list rules of the separation rounds (madam_amd/csrc/madam_cutlists.h,
host C++), checked on the CPU through tests/host/cutlists_dump.cpp:

  top-K list of SA results (cutting_planes.c:247-303, 444-501)
      runs are taken in order; a run is a cut when its cost is below 1;
      while the list is short it is appended (tracking its largest cost),
      then a cost below that largest replaces its entry in place; `best` is
      the smallest cost of all runs.  The reference then looks for the new
      largest entry starting from BIG_NUMBER and keeping a larger one
      (cutting_planes.c:286-296), which never happens: after the first
      replacement every later violated run replaces that same entry.  The
      library keeps this behaviour (node-count parity), and so do these
      tests.
  purge / duplicate check / append (cutting_planes.c:132-214, 316-409,
  512-605)
      cuts whose dual is below 1e-5 are dropped, the rest keep their dual
      and order; a candidate equal (type and vertex tuple) to a survivor is
      skipped, the others are appended with dual 0 -- checked against the
      survivors only, so a purged cut may come back and a candidate listed
      twice is appended twice.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = shutil.which("g++") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cutlists") / "cutlists_dump")
    subprocess.run([CXX, "-O1", "-std=c++17", os.path.join(ROOT, "tests", "host", "cutlists_dump.cpp"),
                    "-o", exe], check=True, capture_output=True)
    return exe


def run_list(exe, k, K, runs):
    lines = [f"list {k}", f"{K} {len(runs)}"]
    lines += [f"{c!r} {t} " + " ".join(map(str, p)) for c, t, p in runs]
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout.split("\n")
    best = float(out[0].split()[1])
    cnt = int(out[1])
    lst = []
    for line in out[2:2 + cnt]:
        f = line.split()
        lst.append((int(f[0]), float(f[1]), tuple(int(v) for v in f[2:])))
    return best, lst


def run_purge(exe, k, cuts, cands):
    lines = [f"purge {k}", str(len(cuts))]
    lines += [f"{t} {d!r} " + " ".join(map(str, p)) for t, p, d in cuts]
    lines += [str(len(cands))]
    lines += [f"{t} " + " ".join(map(str, p)) for t, p in cands]
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout.split("\n")
    cnt = int(out[0])
    res = []
    for line in out[1:1 + cnt]:
        f = line.split()
        res.append((int(f[0]), tuple(int(v) for v in f[2:]), float(f[1])))
    return res


def test_top_k_known_order(harness):
    runs = [(0.9, 1, (0, 1, 2, 3, 4)), (1.2, 2, (5, 6, 7, 8, 9)), (0.8, 2, (1, 2, 3, 4, 5)),
            (0.7, 3, (2, 3, 4, 5, 6)), (0.85, 1, (3, 4, 5, 6, 7))]
    best, lst = run_list(harness, 5, 2, runs)
    assert best == 0.7
    # fill [0.9, 0.8] (largest: entry 0); 0.7 < 0.9 replaces entry 0; the
    # rescan finds no "largest", so 0.85 replaces entry 0 again
    assert [(t, v) for t, v, _ in lst] == [(1, 0.85), (2, 0.8)]
    assert lst[0][2] == (3, 4, 5, 6, 7)


@pytest.mark.parametrize("k,K,R,seed", [(5, 10, 200, 1), (7, 25, 300, 2), (5, 400, 100, 3)])
def test_top_k_fill_then_one_entry_replaced(harness, k, K, R, seed):
    rng = np.random.default_rng(seed)
    costs = rng.permutation(np.linspace(0.2, 1.6, R))  # distinct
    runs = [(float(c), int(rng.integers(1, 4)), tuple(int(v) for v in rng.choice(50, k, replace=False)))
            for c in costs]
    best, lst = run_list(harness, k, K, runs)
    assert best == costs.min()
    viol = [r for r in runs if r[0] < 1.0]
    got = [(v, t, p) for t, v, p in lst]
    if len(viol) <= K:
        assert got == [(c, t, p) for c, t, p in viol]
        return
    fill = viol[:K]
    s = max(range(K), key=lambda q: (fill[q][0], -q))  # first entry of the largest cost
    later = viol[K:]
    first = next((q for q, r in enumerate(later) if r[0] < fill[s][0]), None)
    want = list(fill)
    if first is not None:
        want[s] = later[-1]  # from the first replacement on, every violated run lands there
    assert got == [(c, t, p) for c, t, p in want]


@pytest.mark.parametrize("k", [5, 7])
def test_purge_and_append(harness, k):
    A, B, C, D, E = [(1 + q % 3, tuple(range(q, q + k))) for q in range(5)]
    cuts = [(*A, 0.1), (*B, 0.0), (*C, 1e-6), (*D, 1e-5)]
    cands = [A, E, E, B]
    got = run_purge(harness, k, cuts, cands)
    assert got == [(*A, 0.1), (*D, 1e-5), (*E, 0.0), (*E, 0.0), (*B, 0.0)]
    # same type and tuple is a duplicate; another type on the same tuple is not
    other = (A[0] % 3 + 1, A[1])
    got = run_purge(harness, k, [(*A, 0.5)], [other, A])
    assert got == [(*A, 0.5), (*other, 0.0)]
