"""bench.py's command-line contract that holds without a GPU: one GPU only
(--gpus N > 1 exits with a message, before touching a device), and every
workload name is accepted by the parser."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_workloads_listed_in_help():
    r = run("--help")
    assert r.returncode == 0
    for w in ("c2-root", "c3-batch", "c1-bab", "bab150", "ns-batch", "c4-learned", "c5"):
        assert w in r.stdout
