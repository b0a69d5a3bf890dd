"""bench.py's command line without a GPU: more than one GPU is refused with a
message before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_steps_must_be_whole_chunks():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "150"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "multiples of --chunk" in r.stderr
