"""Drop-in check: MADAM's own program (main.c, bab_functions.c, heap.c,
evaluate.c, process_input.c, unchanged) linked with
madam_amd/dropin/madam_sdpbound.c in place of bounding.c and admm.c, so that
every SDPbound call of its B&B runs on the GPU through libmadam_bound.  The
binary is oracle/_ref/madam_dropin, built by `make -C oracle dropin` where
the reference tree exists.  Run with MPI as the reference runs (master +
one worker) on g05_60.0 with the parameter values of the reference's params
file, it must print the proven optimum 536 (both projection forms).
"""
import os
import re
import shutil
import subprocess

import pytest

from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "madam_dropin")
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


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/madam_dropin not built")
@pytest.mark.skipif(not os.path.exists(MPIEXEC), reason="mpiexec absent")
@pytest.mark.parametrize("projection", ["0", "1"])
def test_reference_bab_with_gpu_sdpbound(tmp_path, projection):
    shutil.copy(os.path.join(ROOT, "tests", "golden", "rudy", "g05_60.0"), tmp_path / "g05_60.0")
    write_params(tmp_path / "params")
    env = dict(os.environ, MADAM_PROJECTION=projection, MKL_NUM_THREADS="1")
    print("running", BIN, "projection", projection, flush=True)
    r = subprocess.run([MPIEXEC, "-np", "2", BIN, "g05_60.0", "params"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=150)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    out = (tmp_path / "g05_60.0.output").read_text()
    m = re.search(r"Maximum value = (\d+)", out)
    assert m and int(m.group(1)) == 536
    rb = float(re.search(r"Root node bound = ([0-9.]+)", out).group(1))
    assert rb >= 536.0
