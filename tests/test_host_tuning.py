"""The sign-iteration layout table (madam_load_tuning / madam_tuned_layout,
host only): lookup of the cell for an order and a number of active nodes,
rejection of malformed tables, and the committed table (written by
madam_amd/tune.py on an MI355X) loading and covering every order."""
import os

import pytest

import madam_amd as G

TABLE = """# comment
16 1 2
16 8 1
512 1 2   0.4 0.3 0.5
512 32 3
512 64 1
"""


@pytest.fixture
def restore():
    yield
    G.load_tuning(G.TUNE_TABLE if os.path.exists(G.TUNE_TABLE) else None)


def test_lookup(tmp_path, restore):
    p = tmp_path / "t.tsv"
    p.write_text(TABLE)
    G.load_tuning(p)
    # (order, active nodes) -> layout: the smallest tabulated n_p >= the
    # order rounded up to 16 (else the largest), the largest tabulated node
    # count <= active (else the smallest)
    cases = {(10, 1): 2, (16, 4): 2, (16, 100): 1, (100, 40): 3, (512, 64): 1, (512, 63): 3,
             (600, 0): 2, (1024, 500): 1, (17, 8): 2, (17, 40): 3}
    for (n, a), lay in cases.items():
        assert G.tuned_layout(n, a) == lay, (n, a)
    G.load_tuning(None)
    assert G.tuned_layout(512, 1) == 0


@pytest.mark.parametrize("text,msg", [("16 1 4\n", "layout 1..3"), ("15 1 1\n", "multiple"),
                                      ("abc\n", "expected"), ("# only\n", "no cells"),
                                      ("2048 1 1\n", "up to 1024")])
def test_rejects_bad_tables(tmp_path, restore, text, msg):
    p = tmp_path / "bad.tsv"
    p.write_text(text)
    with pytest.raises(G.MadamError, match=msg):
        G.load_tuning(p)


def test_missing_file(tmp_path, restore):
    with pytest.raises(G.MadamError, match="cannot open"):
        G.load_tuning(tmp_path / "nope.tsv")


@pytest.mark.skipif(not os.path.exists(G.TUNE_TABLE), reason="no committed table")
def test_committed_table(restore):
    G.load_tuning(G.TUNE_TABLE)
    rows = [ln.split() for ln in open(G.TUNE_TABLE) if ln.strip() and not ln.startswith("#")]
    nps = sorted({int(r[0]) for r in rows})
    assert nps[0] <= 64 and nps[-1] >= 1024
    for r in rows:
        ms = [float(x) for x in r[3:6]]
        assert int(r[2]) == 1 + ms.index(min(ms))  # the fastest layout was recorded
    for n in (30, 60, 100, 150, 225, 512, 1024):
        assert G.tuned_layout(n, 1) in (1, 2, 3)
