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

Both are built by `make -C oracle all` where the reference tree exists.
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

import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINS = {name: os.path.join(ROOT, "oracle", "_ref", name)
        for name in ("madam_dropin", "madam_dropin_admm")}
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
