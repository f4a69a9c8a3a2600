"""Autotuner of the sign-iteration projection's launch layout.

For each order n_p (a multiple of 16) and number of nodes projected
together, times one projection of synthetic nodes in each of the three
layouts (madam_time_projection: 1 one workgroup per node, 2 one launch per
product over 16 x 16 tiles, 3 the same over 2 x 2 blocks of tiles -- all
give the same bits) and writes the fastest to a table that the library
loads (madam_load_tuning; madam_amd loads madam_amd/ns_tune.tsv when it
exists) to pick the layout of every ADMM projection from the largest
order and the active nodes of the batch.

    python -m madam_amd.tune [--out PATH] [--reps R] [--quick]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

NPS = [16, 32, 48, 64, 80, 96, 112, 128, 144, 160, 176, 192, 224, 256, 320, 384, 448, 512, 640,
       768, 896, 1024]
NBS = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512]


def cells(nps, nbs, max_elems=1 << 27):
    """(n_p, nodes) pairs to time, skipping batches beyond max_elems matrix
    entries (their memory and minutes buy nothing: the one-workgroup layout
    wins long before)."""
    for np_ in nps:
        for nb in nbs:
            if nb == 1 or nb * np_ * np_ <= max_elems:
                yield np_, nb


def tune(out, nps=NPS, nbs=NBS, reps=3, log=print):
    import madam_amd as G
    G.open_device(0)
    rows = []
    t0 = time.time()
    for np_, nb in cells(nps, nbs):
        ms = [G.time_projection(np_, nb, lay, reps) for lay in (1, 2, 3)]
        best = 1 + min(range(3), key=lambda i: ms[i])
        rows.append((np_, nb, best, ms))
        log(f"n_p {np_:4d} nodes {nb:3d}: one-WG {ms[0]:8.3f} ms  tiles {ms[1]:8.3f} ms  "
            f"2x2 tiles {ms[2]:8.3f} ms -> layout {best}  ({time.time() - t0:.0f} s)")
    tmp = out + ".tmp"
    with open(tmp, "w") as f:
        f.write("# sign-iteration projection layout per (n_p, active nodes), written by "
                "madam_amd/tune.py\n")
        f.write("# layout: 1 one workgroup per node, 2 tiles, 3 2x2 tiles; "
                "then ms per projection of the batch in layouts 1 2 3\n")
        f.write("# n_p nodes layout ms1 ms2 ms3\n")
        for np_, nb, best, ms in rows:
            f.write(f"{np_} {nb} {best} {ms[0]:.4f} {ms[1]:.4f} {ms[2]:.4f}\n")
    os.replace(tmp, out)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "ns_tune.tsv"))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--quick", action="store_true", help="a few cells only (smoke run)")
    a = ap.parse_args(argv)
    if a.quick:
        tune(a.out, nps=[64, 512], nbs=[1, 8], reps=1)
    else:
        tune(a.out, reps=a.reps)
    print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
