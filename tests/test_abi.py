"""CPU checks of the C ABI: the library is built and exports every entry
point that include/madam_bound.h declares (no GPU calls)."""
import ctypes as C
import os

import pytest

import madam_amd


def test_library_built_and_exports_every_declared_symbol():
    assert os.path.exists(madam_amd.LIB_PATH), "run __graft_entry__.build() first"
    lib = C.CDLL(madam_amd.LIB_PATH)
    names = madam_amd.exported_symbols()
    assert "madam_admm_batch" in names and "madam_open" in names
    for name in names:
        assert hasattr(lib, name), name


def test_struct_layouts_match_reference_types():
    # admm.h:93-117 and 57-89
    assert C.sizeof(madam_amd.Tri) == 32
    assert C.sizeof(madam_amd.Pent) == 40
    assert C.sizeof(madam_amd.Hept) == 48
    from oracle import oracle_py as O
    assert C.sizeof(O.Params) == 128  # sizeof(Parameters) with gcc on x86-64


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(madam_amd, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(madam_amd, "_lib", None)
    with pytest.raises(madam_amd.MadamError):
        madam_amd.lib()


def test_python_mirror_matches_header_layouts(tmp_path):
    """sizeof and field offsets of the header's structs (compiled with gcc)
    equal those of the ctypes mirror in madam_amd."""
    import subprocess
    src = tmp_path / "sizes.c"
    checks = {
        "madam_bab_ctx": (madam_amd.BabCtx, ["params", "LB", "best_x", "diff", "seed", "projection",
                                             "batch", "nodes", "root_bound", "cg_rtol", "cg_iters",
                                             "ns_fallbacks", "verbose", "ns_layout", "own_incumbent"]),
        "madam_opts": (madam_amd.Opts, ["cg_rtol", "cg_maxit", "projection", "ns_layout"]),
        "madam_admm_job": (madam_amd.AdmmJob, ["n", "L", "tri", "sigma", "tol", "itermax", "nbit",
                                               "f", "Z", "M", "cg_iters", "sweeps"]),
        "madam_node": (madam_amd.Node, ["xfixed", "solX", "fracsol", "level", "upper_bound"]),
        "madam_params": (madam_amd.Params, ["sigma0", "TriIneq", "branchingStrategy"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "madam_bound.h"', "int main(void) {"]
    for sname, (_, fields) in checks.items():
        lines.append(f'printf("{sname} sizeof %zu\\n", sizeof({sname}));')
        for f in fields:
            lines.append(f'printf("{sname} {f} %zu\\n", offsetof({sname}, {f}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    inc = os.path.join(os.path.dirname(madam_amd.LIB_PATH), "..", "..", "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for sname, (cls, fields) in checks.items():
        assert got[(sname, "sizeof")] == C.sizeof(cls), sname
        for f in fields:
            assert got[(sname, f)] == getattr(cls, f).offset, (sname, f)
