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
