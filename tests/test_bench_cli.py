"""bench.py's command line, without a GPU: it measures one GPU, so any other
--gpus value ends with a message and status 2 before anything touches a
device; the options it documents parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused_with_a_message():
    res = subprocess.run([sys.executable, BENCH, "--gpus", "2", "--steps", "1", "--warmup", "0"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2, res
    assert "one GPU" in res.stderr, res.stderr
    assert res.stdout == ""


def test_help_lists_the_legs():
    res = subprocess.run([sys.executable, BENCH, "--help"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    for opt in ("--gpus", "--steps", "--warmup", "--no-cpu", "--no-c3", "--no-periodic",
                "--no-bar", "--no-long", "--kinds"):
        assert opt in res.stdout, opt
