"""bench.py's command line on a host without a GPU: more than one GPU is
refused with a message and exit status 2 before anything touches a device;
--help lists the measurement modes."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                           *args], capture_output=True, text=True, timeout=120,
                          cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_modes():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--forms", "--full-call", "--ref-check", "--batch", "--tune",
                "--cpu-check",
                "--copy-calibration", "--cpu"):
        assert opt in r.stdout
