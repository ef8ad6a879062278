"""bench.py's command line on the host: more than one GPU is refused with a JSON message
before anything touches a device, and the options parse."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_options_parse():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--reference-only", "--reference-solvers", "--eager",
                "--profile-kernel", "--layout", "--no-reference"):
        assert opt in r.stdout
