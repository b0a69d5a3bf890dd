"""The serving workload on the GPU, checked against the oracle.

Requests come from the generators the bench serves (`serving_library`:
2,048 template rods of 32-512 elements, log-uniform, with the ragged batch's
materials, tip forces and each rod's own stable dt; `serving_requests`: 1-64
rods per request), with horizons cut to 1,000-2,000 steps so the test fits the
suite.  Requests are admitted in waves while earlier ones run.  For a sample
of completed requests every rod is compared with the oracle stepping the
same rod alone at its own dt (FP64 server: bitwise against the oracle built
with the device library's sin / cos / acos, and within 1e-7 of the
C-library oracle, the bound the spread of those functions' last bits keeps
on the 1024 x 64 batch over 10,000 steps (tests/test_oracle_conditioning.py);
FP32-bulk: positions to 1e-6 of the rod length), and with the same
request served alone (bitwise).  The nearest
reference analogue of admission is first-fit blocking of rods at emplace
time (elastica/Simulator/Configuration/ModuleConfiguration/BlockingPolicy.hpp:202-222).
"""
import os

import numpy as np
import pytest

from elasticapp_amd import FP32_BULK, FP64, Server
from elasticapp_amd.workloads import Workload, build_block, serving_library, serving_requests
from tests.gpu_helpers import STATE, oracle_system

pytestmark = pytest.mark.gpu

CLASSES = [(64, 256), (128, 256), (256, 256), (512, 256)]
N_REQUESTS = 50


@pytest.fixture(scope="module")
def library():
    wl, dts = serving_library()
    lib = build_block(wl)
    lib.set_rod_dt(dts)
    yield wl, dts, lib
    lib.close()


def _oracle(wl, dts, rods, steps, variant=None):
    """The request's rods, each in a block of its own, stepped `steps` steps
    at its own dt by the oracle."""
    sub = Workload("sub", [wl.n_elems[r] for r in rods], [wl.states[r] for r in rods], wl.dt, 0,
                   forces=[(i,) + tuple(wl.forces[r][1:]) for i, r in enumerate(rods)])
    sysm = oracle_system(sub, max_block_elems=0, variant=variant)
    sysm.step_per_block_dt(np.array([dts[r] for r in rods]), steps, min(16, len(os.sched_getaffinity(0))))
    return sysm


def _serve(lib, reqs, precision, waves=((0, 20), (20, 35), (35, N_REQUESTS))):
    """Serve the requests, submitting each wave after the previous one has
    run for two ticks; returns every result read back and the tick count."""
    srv = Server(lib, CLASSES, precision=precision, steps_per_tick=500, hbm_budget_bytes=200e9)
    results, ticks, wave = {}, 0, 0
    while len(results) < len(reqs):
        if wave < len(waves) and ticks == 2 * wave:
            for rid in range(*waves[wave]):
                srv.submit(rid, reqs[rid][0], reqs[rid][1])
            wave += 1
        for rid in srv.tick():
            results[rid] = [{t: srv.read(rid, k, t) for t in STATE} for k in range(len(reqs[rid][0]))]
            srv.release(rid)
        ticks += 1
        assert ticks < 200
    st = srv.status()
    assert st["queued"] == 0 and st["running"] == 0
    srv.close()
    return results, ticks


@pytest.mark.parametrize("precision", [FP64, FP32_BULK])
def test_served_requests_match_the_oracle(library, precision):
    wl, dts, lib = library
    reqs = serving_requests(N_REQUESTS, len(wl.n_elems), seed=3104, min_steps=1000, max_steps=2000)
    # more rods than slots of a class at once: the later waves wait for slots
    assert sum(len(r) for r, _ in reqs) > CLASSES[0][1]
    results, ticks = _serve(lib, reqs, precision)
    print("served %d requests (%d rods) in %d ticks" % (len(reqs), sum(len(r) for r, _ in reqs), ticks), flush=True)
    for rid, (rods, steps) in enumerate(reqs):
        for k in range(len(rods)):
            for t in STATE:
                assert np.isfinite(results[rid][k][t]).all(), (rid, k, t)
    # a sample of requests, rod by rod against the oracle
    sample = [0, 13, 27, 41, 49]
    worst = {}
    for rid in sample:
        rods, steps = reqs[rid]
        sysm = _oracle(wl, dts, list(rods), steps)
        dev = _oracle(wl, dts, list(rods), steps, variant="devmath") if precision == FP64 else None
        for k, r in enumerate(rods):
            L = float(wl.states[r]["ReferenceElementLength"].sum())
            for t in STATE:
                c = sysm.rod_field(k, t)
                g = results[rid][k][t]
                if precision == FP64:
                    assert np.array_equal(g, dev.rod_field(k, t)), (rid, k, t)
                    e = float(np.abs(g - c).max() / max(np.abs(c).max(), 1e-300))
                    assert e <= 1e-7, (rid, k, t, e)
                elif t == "Position":
                    e = float(np.abs(g - c).max() / L)
                    assert e <= 1e-6, (rid, k, t, e)
                else:
                    continue
                worst[t] = max(worst.get(t, 0.0), e)
        sysm.close()
        if dev is not None:
            dev.close()
    print("worst error against the C-library oracle over the sample:", worst, flush=True)


def test_served_request_alone_equals_in_the_mix(library):
    """Requests served in the mix equal the same requests served alone,
    bitwise (FP32-bulk, the bench's precision)."""
    wl, dts, lib = library
    reqs = serving_requests(N_REQUESTS, len(wl.n_elems), seed=3104, min_steps=1000, max_steps=2000)
    mix, _ = _serve(lib, reqs, FP32_BULK)
    for rid in (3, 30, 44):
        alone, _ = _serve(lib, [reqs[rid]], FP32_BULK, waves=((0, 1),))
        for k in range(len(reqs[rid][0])):
            for t in STATE:
                assert np.array_equal(alone[0][k][t], mix[rid][k][t]), (rid, k, t)
