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


def test_committed_counter_summary_feeds_the_rooflines():
    """The counter summary under profiles/ holds the lines bench.py reads for
    the bench tile shape: HBM bytes per launch (100 steps and 1 step) and the
    VALU instructions per wave and step, with plausible values."""
    sys.path.insert(0, ROOT)
    import bench
    t100, t1 = bench.measured_traffic(1024, 2, 100), bench.measured_traffic(1024, 2, 1)
    assert t100 is not None and t1 is not None
    # one pass over state and constants: 144 B of FP64 state read and written
    # per element plus FP32 constants, for the 994,351 elements of the batch
    assert 0.3e9 < t100 < 0.5e9 and 0.3e9 < t1 < 0.5e9
    # the contact batch: 12.8 M elements, one pass over state and constants per launch
    tc = bench.measured_traffic(512, 2, 100, tag="traffic_contact", path=bench.CONTACT_TRAFFIC_FILE)
    assert tc is not None and 4e9 < tc < 6e9
    waves, ipws = bench.measured_valu(1024, 2, 100)
    assert waves % 8 == 0 and 500 < ipws < 1200
    assert bench.measured_valu(512, 1, 100) is None
