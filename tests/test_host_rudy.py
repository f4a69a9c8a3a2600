"""The library's rudy reader (madam_read_rudy, readData of
process_input.c:182-312) against the oracle's construction of SP->L from
the same edges (its restatement of process_input.c:235-299): bit for bit
on the golden instances and on a file with fractional weights and a
repeated edge; and the reference's reading errors as MADAM_E_ARG with a
message.  Host only: no GPU calls."""
import os

import numpy as np
import pytest

import madam_amd as G
from oracle import oracle_py as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FILES = [os.path.join(GOLD, "rudy", f) for f in ("g05_60.0", "g05_100.0", "w01_100.3")] + \
        [os.path.join(GOLD, "rudy150", "pw05_150.0")]


@pytest.mark.parametrize("path", FILES, ids=os.path.basename)
def test_reader_matches_oracle_bit_for_bit(path):
    got = G.read_rudy(path)
    ref = O.laplacian_from_edges(*O.read_rudy(path))
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_fractional_weights_and_repeated_edge(tmp_path):
    rng = np.random.default_rng(4)
    n = 23
    lines = []
    for _ in range(80):
        i, j = rng.choice(n, size=2, replace=False) + 1
        lines.append(f"{i} {j} {rng.uniform(-3, 3):.17g}")
    lines.append(lines[5].rsplit(" ", 1)[0] + " 0.125")  # overwrites an earlier edge
    p = tmp_path / "g.rudy"
    p.write_text(f"{n} {len(lines)}\n" + "\n".join(lines) + "\n")
    got = G.read_rudy(p)
    nv, ei, ej, w = O.read_rudy(str(p))
    ref = O.laplacian_from_edges(nv, ei, ej, w)
    assert np.array_equal(got, ref)
    # the construction's identities: symmetric, last row = row sums of the
    # Laplacian block, corner = their total
    assert np.array_equal(got, got.T)
    assert np.allclose(got[-1, :-1], got[:-1, :-1].sum(axis=1))


@pytest.mark.parametrize("text,msg", [
    ("", "bad header"),
    ("0 0\n", "bad header"),
    ("5 2\n1 2 1\n", "edge 2 unreadable"),
    ("5 1\n1 6 1\n", "outside 1..5"),
    ("5 1\n0 2 1\n", "outside 1..5"),
    ("5 x\n", "bad header"),
])
def test_reading_errors(tmp_path, text, msg):
    p = tmp_path / "bad.rudy"
    p.write_text(text)
    with pytest.raises(G.MadamError, match=msg):
        G.read_rudy(p)


def test_missing_file_and_capacity(tmp_path):
    with pytest.raises(G.MadamError, match="cannot open"):
        G.read_rudy(tmp_path / "nope")
    with pytest.raises(G.MadamError, match="vertices"):
        G.read_rudy(os.path.join(GOLD, "rudy", "g05_60.0"), capN=40)
