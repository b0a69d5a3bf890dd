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


def test_twenty_steps_after_five_warmup_pass_the_argument_checks():
    """`--steps 20 --warmup 5` (fewer steps than a chunk, not whole chunks)
    gets past the argument checks: without a GPU the run stops later, at the
    device, not with the usage exit code 2 or an argument message."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "1", "--steps", "20", "--warmup", "5"],
                       capture_output=True, text=True, timeout=300, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert r.returncode != 2, r.stderr[-2000:]
    assert "must be" not in r.stderr and "multiples" not in r.stderr
    assert "workload ragged_4096" not in r.stderr or "20 timed steps after 5 warm-up steps, 20 steps per launch" in r.stderr


def test_step_counts_that_cannot_run_are_refused():
    for argv in (["--steps", "0"], ["--warmup", "-1"], ["--chunk", "0"]):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + argv, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 2 and "must be" in r.stderr, argv


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
    # a launch of another step count takes the nearest one measured for the shape
    w2, i2, at = bench.measured_valu(1024, 2, 93, exact=False)
    assert at in (20, 100) and w2 == waves
    b, at = bench.measured_traffic(1024, 2, 93, exact=False)
    assert at in (20, 100) and 0.3e9 < b < 0.5e9
