"""CPU checks of bench.py's command-line contract that need no GPU: more than one GPU is refused
with a message (this build measures one GPU's share), before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode != 0
    assert "one GPU" in res.stderr
    assert res.stdout.strip() == ""  # no JSON line


def test_help_lists_the_workload_switches():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--config", "--dtype", "--precision"):
        assert flag in res.stdout
