"""bench.py's command line on the CPU: more than one GPU is refused with a
message and exit status 2 (the build measures one GPU), before any device
work; --help lists the options the driver and the profiles use."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_options():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-breakdown", "--profile-pass", "--only"):
        assert opt in r.stdout, opt
