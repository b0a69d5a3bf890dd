"""The small configurations run in full against the CPU oracle.

  * Timoshenko cantilever (1 rod x 100 elements, clamped, tip force, damping
    0.1, dt 3e-4): 50,000 fused FP64 steps and 2,000 op-by-op steps against
    the oracle, and the FP32-bulk tip within 1e-6 of the rod length;
  * 1024 rods x 64 elements with random rest curvature and velocities:
    after 10,000 fused FP64 steps positions and directors within 1e-10 of
    the oracle, velocities and angular velocities within 1e-7 of their
    (decayed) size; 500 op-by-op steps within 1e-12 of 500 fused FP64 steps.  The velocity bound is
    the batch's own conditioning: two builds of the oracle that differ only
    in the last bits of sin / cos / acos spread by 3.4e-8 and 5.9e-8 in these
    velocities after 10,000 steps (tests/test_oracle_conditioning.py), and
    against the oracle built with the kernels' own sin / cos / acos the fused
    path is bitwise equal (tests/test_gpu_device_math.py).
"""
import os

import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64
from elasticapp_amd.workloads import build_block, timoshenko_cantilever, uniform_1024x64
from tests.gpu_helpers import STATE, oracle_system, rel_state_error

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def test_timoshenko_against_oracle():
    wl = timoshenko_cantilever(steps=50000)
    fused = build_block(wl)
    fused.step_fused(wl.steps, FP64)
    sysm = oracle_system(wl, max_block_elems=1024)
    sysm.step(wl.dt, wl.steps, THREADS)
    err = rel_state_error(fused, sysm)
    assert max(err.values()) <= 1e-10, err
    f32 = build_block(wl)
    f32.step_fused(wl.steps, FP32_BULK)
    tip32, tip = f32.get_rod("Position", 0)[:, -1], sysm.rod_field(0, "Position")[:, -1]
    assert np.abs(tip32 - tip).max() <= 1e-6 * wl.info["length"], (tip32, tip)
    # the op-by-op path over the first 2,000 steps
    ops = build_block(wl)
    ops.step(wl.dt, 2000)
    s2 = oracle_system(wl, max_block_elems=1024)
    s2.step(wl.dt, 2000, THREADS)
    err = rel_state_error(ops, s2)
    assert max(err.values()) <= 1e-12, err


def test_uniform_batch_against_oracle():
    wl = uniform_1024x64(steps=10000)
    fused = build_block(wl)
    fused.step_fused(wl.steps, FP64)
    assert not fused.check_finite().any()
    sysm = oracle_system(wl, max_block_elems=1024)
    sysm.step(wl.dt, wl.steps, THREADS)
    err = rel_state_error(fused, sysm)
    assert err["Position"] <= 1e-10 and err["Director"] <= 1e-10, err
    # the C-library oracle: velocities within the spread that two builds of
    # the oracle differing only in the last bits of sin / cos / acos show on
    # this batch after 10,000 steps (3.4e-8 and 5.9e-8,
    # tests/test_oracle_conditioning.py); against the oracle built with the
    # device library's functions the fused path is bitwise equal
    # (tests/test_gpu_device_math.py)
    assert err["Velocity"] <= 1e-7 and err["AngularVelocity"] <= 1e-7, err
    # the op-by-op path against the fused one on the device (same math
    # library, so the conditioning above does not enter)
    ops = build_block(wl)
    ops.step(wl.dt, 500)
    f500 = build_block(wl)
    f500.step_fused(500, FP64)
    for r in range(0, 1024, 97):
        for t in STATE:
            a, b = ops.get_rod(t, r), f500.get_rod(t, r)
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), (r, t)


def _driver(*args):
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "elasticapp_amd", "timoshenko")
    out = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=120, check=True).stdout
    w = out.split()
    return np.array([float(v) for v in w[w.index("tip") + 1:w.index("tip") + 4]])


def test_cpp_driver_matches_oracle(tmp_path):
    """The C++ host layer (erod.hpp): the Timoshenko driver's tip after 2,000
    steps on the op-by-op and fused paths, and resumed from a checkpoint
    halfway, against the oracle on the same configuration."""
    wl = timoshenko_cantilever(steps=2000)
    sysm = oracle_system(wl, max_block_elems=1024)
    sysm.step(wl.dt, wl.steps)
    tip = sysm.rod_field(0, "Position")[:, -1]
    fused, plain = _driver(2000), _driver(2000, "--plain")
    resumed = _driver(2000, "--resume", tmp_path / "t.ckpt")
    assert np.array_equal(fused, resumed)
    for got in (fused, plain):
        assert np.abs(got - tip).max() <= 1e-12 * np.abs(tip).max(), (got, tip)
