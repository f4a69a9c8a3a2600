"""Build libmadam_bound.so for gfx950 with hipcc (in-tree, no JIT cache)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libmadam_bound.so")
SOURCES = ["madam_device.hip", "madam_sep.hip", "madam_ipm.hip", "madam_gw.hip", "madam_ns.hip",
           "madam_host.hip", "madam_bab.hip", "madam_mlp.hip",
           "madam_cert.hip", "madam_noalloc.hip", "madam_subp.hip", "madam_ipmt.hip",
           "madam_bjac.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed-FP32 instructions (v_pk_fma_f32 and kin): some of their forms
    # have given silently wrong results on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = [os.path.join(LIBDIR, src.replace(".hip", ".o")) for src in SOURCES]
    cmds = [[HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
            for src, obj in zip(SOURCES, objs)]
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    # the translation units compile independently: a few at a time
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(cmds), (os.cpu_count() or 1), 8))
    with ThreadPoolExecutor(jobs) as ex:
        for _ in ex.map(lambda c: subprocess.run(c, check=True), cmds):
            pass
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
