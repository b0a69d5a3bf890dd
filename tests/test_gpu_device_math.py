"""The FP64 device paths against the oracle built with the device library's
sin, cos and acos (oracle/liboracle_devmath.so): with the transcendental
functions equal on both sides, the op-by-op and fused FP64 steps compute the
restatement's arithmetic exactly, so the comparison is bitwise, over the
configurations' full horizons.

Against the oracle with the C library's functions (the reference's), the
same runs differ by what the trajectories make of last-bit differences of
acos / sin / cos: on the 1024 x 64 batch after 10,000 steps about 3e-8 of
the (decayed) velocities, the spread two builds of the oracle itself show
(tests/test_oracle_conditioning.py).
"""
import os

import numpy as np
import pytest

from elasticapp_amd import FP64
from elasticapp_amd.workloads import build_block, rods_on_plane, timoshenko_cantilever, uniform_1024x64
from tests.gpu_helpers import STATE, oracle_system, random_workload, rel_state_error

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _bitwise(blk, sysm, rods=None):
    """Every state value of every rod equal; on a mismatch the message names
    the worst relative difference per tag."""
    bad = {}
    for r in (range(len(blk.n_elems)) if rods is None else rods):
        for t in STATE:
            g, c = blk.get_rod(t, r), sysm.rod_field(r, t)
            if not np.array_equal(g, c):
                e = float(np.abs(g - c).max() / max(np.abs(c).max(), 1e-300))
                bad[t] = max(bad.get(t, 0.0), e)
    return bad


def test_ragged_batch_with_hooks_bitwise():
    wl = random_workload(40, 2, 200, seed=91, steps=300)
    sysm = oracle_system(wl, variant="devmath")
    sysm.step(wl.dt, wl.steps)
    ops = build_block(wl)
    ops.step(wl.dt, wl.steps)
    fused = build_block(wl)
    fused.step_fused(wl.steps, FP64)
    assert _bitwise(ops, sysm) == {}
    assert _bitwise(fused, sysm) == {}


def test_timoshenko_cantilever_bitwise():
    wl = timoshenko_cantilever(steps=50000)
    sysm = oracle_system(wl, max_block_elems=1024, variant="devmath")
    sysm.step(wl.dt, wl.steps)
    fused = build_block(wl)
    fused.step_fused(wl.steps, FP64)
    assert _bitwise(fused, sysm) == {}


def test_uniform_batch_bitwise():
    """1024 rods x 64 elements, 10,000 steps: the fused FP64 step equals the
    device-math oracle bit for bit, positions, directors and velocities; the
    C-library oracle differs by the conditioning spread (below 1e-7 of the
    velocities, 1e-12 of the positions)."""
    wl = uniform_1024x64(steps=10000)
    fused = build_block(wl)
    fused.step_fused(wl.steps, FP64)
    sysm = oracle_system(wl, max_block_elems=1024, variant="devmath")
    sysm.step(wl.dt, wl.steps, THREADS)
    assert _bitwise(fused, sysm) == {}
    libm = oracle_system(wl, max_block_elems=1024)
    libm.step(wl.dt, wl.steps, THREADS)
    err = rel_state_error(fused, libm)
    print("fused FP64 against the C-library oracle after 10,000 steps:", err, flush=True)
    assert err["Position"] <= 1e-12 and err["Director"] <= 1e-9, err
    assert err["Velocity"] <= 1e-7 and err["AngularVelocity"] <= 1e-7, err


def test_contact_cut_bitwise():
    wl = rods_on_plane(n_rods=48, n=128, steps=300)
    sysm = oracle_system(wl, variant="devmath")
    sysm.step(wl.dt, wl.steps)
    ops = build_block(wl)
    ops.step(wl.dt, wl.steps)
    fused = build_block(wl)
    fused.step_fused(wl.steps, FP64)
    assert _bitwise(ops, sysm) == {}
    assert _bitwise(fused, sysm) == {}
