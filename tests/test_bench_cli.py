"""bench.py's command line on the CPU: it parses, lists every workload and
refuses more than one GPU with a message (the engine runs on one GPU)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_help_lists_workloads():
    r = _run("--help")
    assert r.returncode == 0
    for w in ("plummer10M-h8-p8", "plummer10M-h8-p5", "uniform1M-h7-p7", "batch4096-p5-r8", "compare20k-h5-p7"):
        assert w in r.stdout
    assert "hilo" in r.stdout


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
