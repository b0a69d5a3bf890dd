"""How far the 1024-rod x 64-element batch carries last-bit differences of
acos / sin / cos over its 10,000 steps, measured with the oracle alone (CPU):
three builds of the same restatement that differ only in those functions
(the C library's, as the reference; the device library's; the C library's
with a quarter of its results moved by one ulp).

Measured (this test's bounds hold each with a margin of about 2):
velocities ~3.4e-8 and angular velocities ~5.9e-8 of their size after
10,000 steps, directors ~4e-11, positions ~3e-14.  This is the floor under
any comparison of an FP64 run with the reference's: a different C library
on the reference's own side gives the same spread.  The device FP64 paths
equal the device-library build bit for bit (tests/test_gpu_device_math.py),
so against the C-library build they show exactly this spread.
"""
import os

import numpy as np
import pytest

from elasticapp_amd.workloads import uniform_1024x64
from tests.gpu_helpers import STATE, oracle_system

pytestmark = pytest.mark.slow

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _run(wl, variant):
    s = oracle_system(wl, max_block_elems=1024, variant=variant)
    s.step(wl.dt, wl.steps, THREADS)
    out = [{t: s.rod_field(r, t).copy() for t in STATE} for r in range(len(wl.n_elems))]
    s.close()
    return out


def _spread(a, b):
    out = {}
    for ra, rb in zip(a, b):
        for t in STATE:
            e = float(np.abs(ra[t] - rb[t]).max() / max(np.abs(rb[t]).max(), 1e-300))
            out[t] = max(out.get(t, 0.0), e)
    return out


def test_last_bit_differences_of_the_transcendentals_reach_3e_8_of_the_velocities():
    wl = uniform_1024x64(steps=10000)
    libm, dev, pert = _run(wl, "libm"), _run(wl, "devmath"), _run(wl, "perturbed")
    for name, s in (("device library", _spread(dev, libm)), ("perturbed C library", _spread(pert, libm))):
        print(name, "against the C library:", s)
        assert 1e-8 <= s["Velocity"] <= 1e-7 and 1e-8 <= s["AngularVelocity"] <= 1.2e-7, (name, s)
        assert s["Director"] <= 1e-10 and s["Position"] <= 1e-13, (name, s)
    # in absolute terms: a few 1e-9 m/s against initial velocities N(0, 1e-2) m/s
    dv = max(np.abs(a["Velocity"] - b["Velocity"]).max() for a, b in zip(dev, libm))
    assert 1e-10 < dv < 1e-8
