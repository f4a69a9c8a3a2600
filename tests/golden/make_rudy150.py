"""Copies two n=150 instances and their recorded results from the
reference's Instances/rudy/rudy_150/ro_1 (graph files and the root bound,
optimum and solution printed in the .output files) into tests/golden/.

Run here only (reads /root/reference); the outputs are committed data."""
import json
import os
import shutil

SRC = "/root/reference/Instances/rudy/rudy_150/ro_1"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rudy150")


def parse(path):
    rec = {}
    for line in open(path):
        if line.startswith("Root node bound ="):
            rec["root_bound"] = float(line.split("=")[1])
        elif line.startswith("Maximum value ="):
            rec["optimum"] = float(line.split("=")[1])
        elif line.startswith("Solution ="):
            rec["solution"] = [int(v) for v in line.split("(")[1].split(")")[0].split()]
        elif line.strip().startswith("TriIneq ="):
            rec["TriIneq"] = int(line.split("=")[1])
    return rec


def main():
    os.makedirs(DST, exist_ok=True)
    out = {}
    for name in ("pw05_150.0", "pw05_150.1"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
        out[name] = parse(os.path.join(SRC, name + ".output"))
    with open(os.path.join(DST, "recorded.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
